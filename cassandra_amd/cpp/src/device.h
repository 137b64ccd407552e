// Shared helpers of the C++ drop-in layer: the process-wide device context and
// error translation (C-ABI status -> std::runtime_error, as the reference throws).
#pragma once
#include <cstddef>
#include <stdexcept>
#include <string>

#include "../../../include/cassandra_amd.h"

namespace cad_cpp {

cad_context* Context();                 // device 0, created on first use
void Check(int rc);                     // throws std::runtime_error(cad_last_error())

// Device buffer owned by the drop-in layer.
class DevBuf
{
    void* p = nullptr;

public:
    explicit DevBuf(std::size_t bytes) { Check(cad_malloc(Context(), bytes, &p)); }
    ~DevBuf() { cad_free(Context(), p); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    template <class T> T* as() const { return static_cast<T*>(p); }
};

}  // namespace cad_cpp
