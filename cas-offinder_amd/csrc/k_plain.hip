// Plain path (the reference launch model): finder passes and the per-key comparer.
#include "kcommon.hpp"
#include <rocprim/device/device_scan.hpp>

namespace cof {

// ------------------------------------------------------------------ plain path

struct FinderArgs {
    DevGenome g;
    PatternDesc pd;
    uint64_t slab_lo, slab_hi;   // window starts [slab_lo, slab_hi), slab_lo tile aligned
    uint32_t *tile_cnt;          // count pass: (locus, flag) entries of each tile
    const uint32_t *tile_incl;   // write pass: inclusive scan of tile_cnt
    unsigned long long *cand_slots;  // (locus, strand) candidates, kCountSlots slots
    uint32_t *loci;              // locus - slab_lo
    uint8_t *flags;              // 0 both strands, 1 forward only, 2 reverse only
    uint64_t cap;
};



// PAM masks of the starts of this workgroup's tile of the slab, with
// poison and slab range applied; mine = starts with a PAM on either strand.
__device__ __forceinline__ uint64_t slab_tile_masks(const FinderArgs &a, uint32_t *s_lo, uint32_t *s_hi,
                                                    uint32_t *s_ot, uint32_t (&fwd)[kWordsPerThread],
                                                    uint32_t (&rev)[kWordsPerThread], uint32_t &mine,
                                                    uint32_t &nmine) {
    const int t = threadIdx.x;
    const int64_t w0 = (int64_t)(a.slab_lo >> 5) + (int64_t)blockIdx.x * kTileWords;
    const uint64_t ppos = a.g.tile_ppos[(uint64_t)w0 * 32 / kTilePos];
    stage_tile(a.g, w0, s_lo, s_hi, s_ot);
    __syncthreads();
    const uint64_t t0 = (uint64_t)w0 * 32, t1 = t0 + kTilePos;
    uint32_t keep[kWordsPerThread];
    uint64_t s0[kWordsPerThread];
    finder_masks(a.pd, s_lo, s_hi, s_ot, kWordsPerThread * t + kHaloL, fwd, rev);  // words 4t .. 4t+3
#pragma unroll
    for (int j = 0; j < kWordsPerThread; j++) {
        s0[j] = t0 + 32ull * (kWordsPerThread * t + j);
        keep[j] = range_mask(s0[j], a.slab_lo, a.slab_hi);
    }
    apply_poison(a.g, a.pd.P, t0 / kTilePos, t1, ppos, s0, keep);
    mine = nmine = 0;
#pragma unroll
    for (int j = 0; j < kWordsPerThread; j++) {
        fwd[j] &= keep[j];
        rev[j] &= keep[j];
        mine += __popc(fwd[j] | rev[j]);
        nmine += __popc(fwd[j]) + __popc(rev[j]);
    }
    return t0;
}

// Plain-path finder, pass 1: entries per tile (no atomics on a shared
// address: the tile offsets come from a scan of these counts).
__global__ __launch_bounds__(kThreads) void finder_count_kernel(FinderArgs a) {
    __shared__ uint32_t s_lo[kLdsWords], s_hi[kLdsWords], s_ot[kLdsWords];
    __shared__ uint32_t s_wsum[4], s_wsum2[4];
    uint32_t fwd[kWordsPerThread], rev[kWordsPerThread], mine, nmine;
    (void)slab_tile_masks(a, s_lo, s_hi, s_ot, fwd, rev, mine, nmine);
    uint32_t total, nstrand;
    (void)block_scan(mine, s_wsum, &total);
    (void)block_scan(nmine, s_wsum2, &nstrand);
    if (threadIdx.x == 0) {
        a.tile_cnt[blockIdx.x] = total;
        if (nstrand) atomicAdd(a.cand_slots + (blockIdx.x % kCountSlots) * kSlotStride, (unsigned long long)nstrand);
    }
}

// Pass 2: the same masks, entries written at the tile's scanned offset in
// position order (cas-offinder.cl:68-70 compacts the same (locus, flag)
// pairs in atomic order).
__global__ __launch_bounds__(kThreads) void finder_write_kernel(FinderArgs a) {
    __shared__ uint32_t s_lo[kLdsWords], s_hi[kLdsWords], s_ot[kLdsWords];
    __shared__ uint32_t s_ent[kListCap];
    __shared__ uint32_t s_wsum[4];
    const int t = threadIdx.x;
    uint32_t fwd[kWordsPerThread], rev[kWordsPerThread], mine, nmine;
    const uint64_t t0 = slab_tile_masks(a, s_lo, s_hi, s_ot, fwd, rev, mine, nmine);
    uint32_t total;
    const uint32_t first = block_scan(mine, s_wsum, &total);
    const uint64_t base = blockIdx.x ? a.tile_incl[blockIdx.x - 1] : 0u;
    // entries (tile offset << 2 | flag) staged in LDS, then written out coalesced
    for (uint32_t pass = 0; pass < total; pass += kListCap) {
        const uint32_t cnt = min(total - pass, (uint32_t)kListCap);
        if (first < pass + cnt && first + mine > pass) {
            uint32_t idx = first;
#pragma unroll
            for (int j = 0; j < kWordsPerThread; j++) {
                const uint32_t wi = (uint32_t)(kWordsPerThread * t + j);
                for (uint32_t m = fwd[j] | rev[j]; m; m &= m - 1, idx++) {
                    const uint32_t b = __builtin_ctz(m);
                    if (idx >= pass && idx < pass + cnt) {
                        const uint32_t fb = (fwd[j] >> b) & 1, rb = (rev[j] >> b) & 1;
                        s_ent[idx - pass] = ((wi * 32 + b) << 2) | (fb && rb ? 0u : fb ? 1u : 2u);
                    }
                }
            }
        }
        __syncthreads();
        for (uint32_t i = t; i < cnt; i += kThreads) {
            const uint32_t e = s_ent[i];
            const uint64_t o = base + pass + i;
            if (o < a.cap) {
                a.loci[o] = (uint32_t)(t0 + (e >> 2) - a.slab_lo);
                a.flags[o] = (uint8_t)(e & 3);
            }
        }
        __syncthreads();
    }
}

struct ComparerArgs {
    DevGenome g;
    uint32_t P, cbase, gbase, pad;
    const uint32_t *table;  // key table of one key (one care class of one key for concrete tables)
    uint64_t slab_lo;
    const uint32_t *loci;
    const uint8_t *flags;
    uint64_t n;
    uint64_t *hkey;
    uint16_t *hmm;
    unsigned long long *count;
    uint64_t cap;
};

template <int NW, bool GEN>
__global__ __launch_bounds__(256) void comparer_kernel(ComparerArgs a) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const uint64_t pos = a.slab_lo + a.loci[i];
    const uint32_t fl = a.flags[i];
    const DevGenome &g = a.g;
    auto rd = [&](int64_t wi, uint32_t &l, uint32_t &h, uint32_t &o) {
        const bool in = wi >= 0;  // the planes carry zero padding words past the end
        l = in ? g.lo[wi] : 0u;
        h = in ? g.hi[wi] : 0u;
        o = in ? g.ot[wi] : ~0u;
    };
    for (uint32_t strand = 0; strand < 2; strand++) {
        if (strand == 0 ? !(fl == 0 || fl == 1) : !(fl == 0 || fl == 2)) continue;
        const int64_t bit = strand == 0 ? (int64_t)pos + a.cbase
                                        : (int64_t)pos + (int64_t)a.P - 32 * NW - (int64_t)a.cbase;
        uint32_t l[1][NW], h[1][NW], o[1][NW];
        window_planes<NW>(rd, bit, strand, l[0], h[0], o[0]);
        const HitSink sink[1] = {HitSink{a.hkey, a.hmm, a.count, a.cap, pos, strand, a.gbase}};
        const bool valid[1] = {true};
        score_table<NW, GEN, 1>((kptr_t)a.table, 1, l, h, o, sink, valid);
    }
}

static FinderArgs finder_args(const DevGenome &g, const PatternDesc &pd, uint64_t slab_lo, uint64_t slab_hi) {
    FinderArgs a{};
    a.g = g;
    a.pd = pd;
    a.slab_lo = slab_lo;
    a.slab_hi = slab_hi;
    return a;
}

uint64_t finder_tiles(uint64_t slab_lo, uint64_t slab_hi) {
    return slab_hi > slab_lo ? (slab_hi - slab_lo + kTilePos - 1) / kTilePos : 0;
}

hipError_t launch_finder_count(const DevGenome &g, const PatternDesc &pd, uint64_t slab_lo, uint64_t slab_hi,
                               uint32_t *tile_cnt, unsigned long long *cand_slots, hipStream_t s) {
    const uint64_t tiles = finder_tiles(slab_lo, slab_hi);
    if (!tiles) return hipSuccess;
    FinderArgs a = finder_args(g, pd, slab_lo, slab_hi);
    a.tile_cnt = tile_cnt;
    a.cand_slots = cand_slots;
    hipLaunchKernelGGL(finder_count_kernel, dim3((uint32_t)tiles), dim3(kThreads), 0, s, a);
    return hipGetLastError();
}

hipError_t scan_tile_counts(void *tmp, size_t *tmp_bytes, const uint32_t *cnt, uint32_t *incl, uint64_t n,
                            hipStream_t s) {
    return rocprim::inclusive_scan(tmp, *tmp_bytes, cnt, incl, (size_t)n, rocprim::plus<uint32_t>(), s);
}

hipError_t launch_finder_write(const DevGenome &g, const PatternDesc &pd, uint64_t slab_lo, uint64_t slab_hi,
                               const uint32_t *tile_incl, uint32_t *loci, uint8_t *flags, uint64_t cap,
                               hipStream_t s) {
    const uint64_t tiles = finder_tiles(slab_lo, slab_hi);
    if (!tiles) return hipSuccess;
    FinderArgs a = finder_args(g, pd, slab_lo, slab_hi);
    a.tile_incl = tile_incl;
    a.loci = loci;
    a.flags = flags;
    a.cap = cap;
    hipLaunchKernelGGL(finder_write_kernel, dim3((uint32_t)tiles), dim3(kThreads), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_comparer(const DevGenome &g, uint32_t P, int nw, bool gen, uint32_t cbase, const uint32_t *table,
                           uint32_t gbase, uint64_t slab_lo, const uint32_t *loci, const uint8_t *flags, uint64_t n,
                           uint64_t *hkey, uint16_t *hmm, unsigned long long *count, uint64_t cap, hipStream_t s) {
    if (n == 0) return hipSuccess;
    ComparerArgs a{g, P, cbase, gbase, 0, table, slab_lo, loci, flags, n, hkey, hmm, count, cap};
    const dim3 grid((uint32_t)((n + 255) / 256)), block(256);
    if (nw == 1) {
        if (gen) hipLaunchKernelGGL((comparer_kernel<1, true>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((comparer_kernel<1, false>), grid, block, 0, s, a);
    } else {
        if (gen) hipLaunchKernelGGL((comparer_kernel<2, true>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((comparer_kernel<2, false>), grid, block, 0, s, a);
    }
    return hipGetLastError();
}

}  // namespace cof
