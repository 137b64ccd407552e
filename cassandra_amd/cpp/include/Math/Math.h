#pragma once
#include "DeviceSolver.h"
#include "MatrixCSR.h"
#include "MatrixProxy.h"
#include "Solver.h"
#include "Vector.h"
