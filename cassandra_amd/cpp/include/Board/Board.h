#pragma once
#include "Position.h"
