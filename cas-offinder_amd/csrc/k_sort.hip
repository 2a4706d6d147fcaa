// Device radix sort of the hit records.
#include "kcommon.hpp"
#include <rocprim/device/device_radix_sort.hpp>

namespace cof {

const char *last_launched_kernel() { return g_kernel_name.c_str(); }
void reset_launched_kernel() { g_kernel_name.clear(); }

// Radix sort of the hit records by their 64-bit key (bits [0, end_bit)).
// With tmp == nullptr only *tmp_bytes is set.
hipError_t sort_hits(void *tmp, size_t *tmp_bytes, const uint64_t *kin, uint64_t *kout, const uint16_t *vin,
                     uint16_t *vout, uint64_t n, unsigned end_bit, hipStream_t s) {
    return rocprim::radix_sort_pairs(tmp, *tmp_bytes, kin, kout, vin, vout, (size_t)n, 0u, end_bit, s);
}

}  // namespace cof
