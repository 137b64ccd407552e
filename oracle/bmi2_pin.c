/* Hardware PEXT/PDEP (x86 BMI2), used only by the CPU tests to pin the
 * oracle's loop-form or_pext/or_pdep (reference: Base/Bit.h:89-98 calls the
 * same instructions).  TEST INFRASTRUCTURE ONLY. */
#include <stdint.h>
#include <immintrin.h>

uint64_t or_hw_pext(uint64_t src, uint64_t mask) { return _pext_u64(src, mask); }
uint64_t or_hw_pdep(uint64_t src, uint64_t mask) { return _pdep_u64(src, mask); }
