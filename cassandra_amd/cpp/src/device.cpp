#include "device.h"

#include <mutex>

namespace cad_cpp {

cad_context* Context()
{
    static std::once_flag once;
    static cad_context* ctx = nullptr;
    static std::string error;
    std::call_once(once, [] {
        if (cad_context_create(0, &ctx) != CAD_OK) {
            error = cad_last_error();
            ctx = nullptr;
        }
    });
    if (!ctx) throw std::runtime_error("no HIP device for the drop-in layer: " + error);
    return ctx;
}

void Check(int rc)
{
    if (rc == CAD_OK) return;
    if (rc == CAD_ERR_SIZE_MISMATCH) throw std::runtime_error("Size mismatch.");
    throw std::runtime_error(cad_last_error());
}

}  // namespace cad_cpp
