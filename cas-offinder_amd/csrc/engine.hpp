// Device context: resident genome, work buffers, stream and the two search
// paths (plain: reference launch model; batched: fused per-query launches).
#pragma once

#include <hip/hip_runtime_api.h>

#include <memory>
#include <string>
#include <vector>

#include "../../include/casoff.h"
#include "devtypes.hpp"
#include "genome.hpp"
#include "query.hpp"

struct cof_query {
    std::shared_ptr<cof::Query> q;
};

struct cof_result {
    std::vector<cof_hit> hits;                      // sorted (query, key, pos, strand)
    std::vector<std::shared_ptr<cof::Query>> queries;
    std::vector<size_t> qbegin;                     // hits of query i: [qbegin[i], qbegin[i+1])
};

namespace cof {

hipError_t launch_finder(const DevGenome &g, const PatternDesc &pd, uint64_t slab_lo, uint64_t slab_hi,
                         uint32_t *loci, uint8_t *flags, unsigned long long *count, uint64_t cap, hipStream_t s);
hipError_t launch_comparer(const DevGenome &g, uint32_t P, int nw, uint32_t cbase, const uint32_t *key,
                           uint32_t key_index, uint32_t qid, uint64_t slab_lo, const uint32_t *loci,
                           const uint8_t *flags, uint64_t n, DevHit *hits, unsigned long long *count, uint64_t cap,
                           hipStream_t s);
hipError_t launch_fused(const DevGenome &g, const PatternDesc &pd, int nw, uint32_t cbase, const uint32_t *keys,
                        uint32_t nkeys, uint32_t qid, DevHit *hits, unsigned long long *count,
                        unsigned long long *cand_count, uint64_t cap, hipStream_t s);

template <typename T>
struct DevBuf {
    T *p = nullptr;
    uint64_t n = 0;  // capacity in elements
    hipError_t reserve(uint64_t want) {
        if (want <= n && p) return hipSuccess;
        if (p) { (void)hipFree(p); p = nullptr; n = 0; }
        hipError_t e = hipMalloc((void **)&p, std::max<uint64_t>(want, 1) * sizeof(T));
        if (e == hipSuccess) n = std::max<uint64_t>(want, 1);
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

}  // namespace cof

struct cof_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    cof::HostGenome genome;
    cof::DevBuf<uint32_t> d_lo, d_hi, d_ot;
    cof::DevBuf<uint64_t> d_poison;
    cof::DevGenome dg{};
    bool genome_ready = false;
    // work buffers
    cof::DevBuf<cof::DevHit> d_hits;
    cof::DevBuf<unsigned long long> d_count;  // [0] hits, [1] finder candidates, [2..] per-query candidates
    cof::DevBuf<uint32_t> d_loci;
    cof::DevBuf<uint8_t> d_flags;
    cof::DevBuf<uint32_t> d_keys;
    cof::DevBuf<uint8_t> d_scratch;
    hipEvent_t ev[8] = {};
    cof_stats stats{};
    uint64_t slab_positions = 1ull << 31;  // plain path window starts per finder launch

    int fail(int code, const std::string &m) {
        err = m;
        return code;
    }
    int hip_fail(hipError_t e, const char *what) {
        err = std::string(what) + ": " + hipGetErrorString(e);
        return COF_E_DEVICE;
    }
    int upload_genome();
    int search(const std::vector<std::shared_ptr<cof::Query>> &qs, int mode, cof_result **out);

private:
    int search_plain(const std::vector<std::shared_ptr<cof::Query>> &qs, uint64_t &nhits);
    int search_batched(const std::vector<std::shared_ptr<cof::Query>> &qs, uint64_t &nhits);
};
