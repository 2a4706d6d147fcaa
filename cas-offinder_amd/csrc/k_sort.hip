// Device radix sort of the hit records, and their packing for the copy to
// the host: a 32-bit window start and one byte mismatches << 1 | strand
// per hit (genomes below 2^32 positions; a count never exceeds the 64
// positions of a pattern), plus per global key index the first hit of the
// key (the hits are sorted by key, so a key's hits are a run).
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

// *dst = the hit counter, ordered after the launches before it on the stream
// (dst may be pinned host memory)
__global__ __launch_bounds__(64) void snap_count_kernel(const unsigned long long *count, unsigned long long *dst) {
    if (threadIdx.x == 0) dst[0] = __hip_atomic_load(count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

hipError_t snap_count(const unsigned long long *count, unsigned long long *dst, hipStream_t s) {
    hipLaunchKernelGGL(snap_count_kernel, dim3(1), dim3(64), 0, s, count, dst);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void fill_u32_kernel(uint32_t *p, uint64_t n, uint32_t v) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

// koff[k] = first hit of key k (keys without hits: the next key's first
// hit; nkeys + 1 entries, preset to n): the hit where the key changes writes
// the keys from the previous hit's key + 1 up to its own.
__global__ __launch_bounds__(256) void pack_hits_kernel(const uint64_t *hkey, const uint16_t *hmm, uint64_t n,
                                                        uint32_t *pos, uint8_t *ms, uint32_t *koff) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint64_t k = hkey[i];
    pos[i] = (uint32_t)(k >> 1);
    ms[i] = (uint8_t)((hmm[i] << 1) | (uint32_t)(k & 1));
    const uint32_t key = (uint32_t)(k >> kHitKeyShift);
    const uint32_t from = i ? (uint32_t)(hkey[i - 1] >> kHitKeyShift) + 1 : 0u;
    for (uint32_t x = from; x <= key; x++) koff[x] = (uint32_t)i;
}

hipError_t pack_hits(const uint64_t *hkey, const uint16_t *hmm, uint64_t n, uint32_t nkeys, uint32_t *pos, uint8_t *ms,
                     uint32_t *koff, hipStream_t s) {
    const uint64_t nk = (uint64_t)nkeys + 1;
    hipLaunchKernelGGL(fill_u32_kernel, dim3((uint32_t)((nk + 255) / 256)), dim3(256), 0, s, koff, nk, (uint32_t)n);
    if (n) hipLaunchKernelGGL(pack_hits_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, hkey, hmm, n, pos, ms, koff);
    return hipGetLastError();
}

}  // namespace cof
