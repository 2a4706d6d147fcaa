// Batched search that allocates nothing (cof_search_noalloc): the key tables
// and launch counters of the search live in a file-scope device array,
// scratch, of kScratchBytes (each GPU the library runs on holds its copy
// until the process ends); the hits go to the caller's device buffers.  The
// host side never names scratch itself: at a context's open on a device
// (noalloc_init, with that device current) hipGetSymbolAddress gives the
// array's device address on that device, kept per device, and every use
// goes through that address.  All calls on one GPU share its array, so calls
// that overlap in time on one GPU race.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

#include "engine.hpp"

namespace cof {

namespace {

constexpr size_t kScratchBytes = 8u << 20;
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch[kMaxDevices];  // device address of scratch on device d (null until noalloc_init)

// zero the counter words (a kernel rather than a memset node: callers may
// capture the search in a HIP graph)
__global__ __launch_bounds__(256) void clear_counters_kernel(unsigned long long *p, size_t n) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

}  // namespace

// Looks up scratch's address on `device` (the current device) once.
hipError_t noalloc_init(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch[device]) return hipSuccess;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch[device] = p;
    return hipSuccess;
}

}  // namespace cof

using namespace cof;

extern "C" int cof_search_noalloc(cof_ctx *c, cof_query *const *queries, int n_queries, uint64_t *hit_keys,
                                  uint16_t *hit_mm, uint64_t cap, uint64_t *n_hits) {
    if (!c || n_queries < 0 || (n_queries && !queries) || !n_hits || (cap && (!hit_keys || !hit_mm)))
        return COF_E_INVALID;
    *n_hits = 0;
    if (c->device < 0) return c->fail(COF_E_UNSUPPORTED, "Host-only context cannot search.");
    if (!c->genome_ready) return c->fail(COF_E_INVALID, "No genome loaded.");
    if (n_queries > 255) return c->fail(COF_E_INVALID, "At most 255 queries per search.");
    if (hipSetDevice(c->device) != hipSuccess) return c->fail(COF_E_DEVICE, "hipSetDevice failed");
    uint8_t *base;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        base = c->device < kMaxDevices ? static_cast<uint8_t *>(g_scratch[c->device]) : nullptr;
    }
    if (!base) return c->fail(COF_E_DEVICE, "The scratch array of this device was not set up.");
    std::vector<std::shared_ptr<Query>> qs;
    for (int i = 0; i < n_queries; i++) {
        if (!queries[i]) return c->fail(COF_E_INVALID, "null query");
        qs.push_back(queries[i]->q);
    }
    // scratch: counter words, then (256-byte aligned) the key tables
    const size_t count_words = search_count_words(qs.size());
    const size_t count_bytes = (count_words * sizeof(unsigned long long) + 255) & ~(size_t)255;
    if (count_bytes >= kScratchBytes)
        return c->fail(COF_E_NOMEM, "Too many queries for the scratch buffer.");
    const size_t nblk = (count_words + 255) / 256;
    hipLaunchKernelGGL(clear_counters_kernel, dim3((uint32_t)nblk), dim3(256), 0, c->lane.stream,
                       reinterpret_cast<unsigned long long *>(base), count_words);
    if (hipGetLastError() != hipSuccess) return c->fail(COF_E_DEVICE, "counter clear launch failed");
    ExtBufs ext{};
    ext.count = reinterpret_cast<unsigned long long *>(base);
    ext.count_words = count_words;
    ext.keys = reinterpret_cast<uint32_t *>(base + count_bytes);
    ext.keys_words = (kScratchBytes - count_bytes) / sizeof(uint32_t);
    ext.hkey = hit_keys;
    ext.hmm = hit_mm;
    ext.cap = cap;
    uint64_t nhits = 0;
    const int rc = run_search_noalloc(c->lane, c->dg, qs, &c->tuner, ext, nhits, c->err);
    *n_hits = nhits;
    return rc;
}
