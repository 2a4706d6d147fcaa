// Batched key tables: fused_kernel (brute force) and fused_seed_kernel (list seed index).
#include "kcommon.hpp"

namespace cof {

struct FusedArgs {
    DevGenome g;
    PatternDesc pd;
    uint32_t cbase, nunits, gbase, tile_stride;
    const uint32_t *table;  // key table (nunits care classes, or nunits general records)
    uint64_t limit;         // valid window starts are [0, limit)
    uint64_t *hkey;
    uint16_t *hmm;
    unsigned long long *count;
    unsigned long long *cand_count;
    uint64_t cap;
};

template <int NW, bool GEN, int BLK>
__global__ __launch_bounds__(kThreads) void fused_kernel(FusedArgs a) {
    __shared__ uint32_t s_lo[kLdsWords], s_hi[kLdsWords], s_ot[kLdsWords];
    __shared__ uint16_t s_list[kListCap];
    __shared__ uint32_t s_wsum[4];
    __shared__ HitStageBuf<kStageCap> s_stage;
    constexpr int W = 1;  // candidates per lane
    const uint32_t P = a.pd.P;
    tile_candidates<W>(a.g, a.pd, a.limit, a.tile_stride, a.cand_count, s_lo, s_hi, s_ot, s_list, s_wsum,
                       stage_ref(s_stage), a.hkey, a.hmm, a.count, a.cap,
                       [&](uint64_t t0, auto rd, const uint32_t (&e)[W], const bool (&valid)[W]) {
                           uint32_t l[W][NW], h[W][NW], o[W][NW];
                           TileSink sink[W];
#pragma unroll
                           for (int w = 0; w < W; w++) {
                               const uint32_t ls = e[w] >> 1, strand = e[w] & 1;
                               const int64_t bit = strand == 0 ? (int64_t)ls + a.cbase + 32 * kHaloL
                                                               : (int64_t)ls + P - 32 * NW - (int64_t)a.cbase +
                                                                     32 * kHaloL;
                               window_planes<NW>(rd, bit, strand, l[w], h[w], o[w]);
                               sink[w] = TileSink{HitSink{a.hkey, a.hmm, a.count, a.cap, t0 + ls, strand, a.gbase},
                                                  stage_ref(s_stage), stage_low(strand, (int32_t)ls)};
                           }
                           score_table<NW, GEN, W, BLK>((kptr_t)a.table, a.nunits, l, h, o, sink, valid);
                       });
}

// ------------------------------------------------------------- seed index path
//
// fused_seed_kernel: a single-class concrete key table with a seed index
// (SeedDesc).  Per candidate, each range's 2-bit codes address the list of
// records whose bases there equal the window's; only those records are
// counted (the same three instructions as fused_kernel), so a candidate costs
// about (keys / 4^range length) counts per range instead of one per key.  A
// window within threshold of a key matches one of the key's ranges exactly,
// so no hit is missed; a hit is emitted from its first exact range only.

struct SeedArgs {
    DevGenome g;
    PatternDesc pd;
    SeedDesc sd;
    uint32_t cbase, tile_stride;
    const uint32_t *table;  // class header [care, n, 0, 0] + records
    const uint32_t *seed;   // slots then entries
    uint64_t limit;
    uint64_t *hkey;
    uint16_t *hmm;
    unsigned long long *count;
    unsigned long long *cand_count;
    uint64_t cap;
};



__global__ __launch_bounds__(kThreads) void fused_seed_kernel(SeedArgs a) {
    __shared__ uint32_t s_lo[kLdsWords], s_hi[kLdsWords], s_ot[kLdsWords];
    __shared__ uint16_t s_list[kSeedListCap];
    __shared__ uint32_t s_wsum[4];
    __shared__ HitStageBuf<kSeedStageCap> s_stage;
    extern __shared__ __align__(16) uint32_t s_seed[];  // records (3 words each), offs, ids
    const SeedDesc &sd = a.sd;
    uint32_t *s_rec = s_seed;
    const uint16_t *s_offs = reinterpret_cast<const uint16_t *>(s_seed + 3 * sd.nrec);
    const uint16_t *s_ids = s_offs + sd.noffs;
    for (uint32_t i = threadIdx.x; i < sd.nrec; i += kThreads) {
        s_rec[3 * i] = a.table[kClassHdr + 4 * i];
        s_rec[3 * i + 1] = a.table[kClassHdr + 4 * i + 1];
        s_rec[3 * i + 2] = a.table[kClassHdr + 4 * i + 2];
    }
    {
        uint32_t *dst = s_seed + 3 * sd.nrec;
        const uint32_t nwords = (sd.noffs + sd.nids + 1) / 2;
        for (uint32_t i = threadIdx.x; i < nwords; i += kThreads) dst[i] = a.seed[i];
    }  // visible after tile_candidates' first barrier
    const uint32_t care = a.table[0];
    const uint32_t P = a.pd.P;
    tile_candidates<1, false, kSeedListCap>(
        a.g, a.pd, a.limit, a.tile_stride, a.cand_count, s_lo, s_hi, s_ot, s_list, s_wsum, stage_ref(s_stage), a.hkey,
        a.hmm, a.count, a.cap, [&](uint64_t t0, auto rd, const uint32_t (&e)[1], const bool (&)[1]) {
            const uint32_t ls = e[0] >> 1, strand = e[0] & 1;
            const int64_t bit = strand == 0 ? (int64_t)ls + a.cbase + 32 * kHaloL
                                            : (int64_t)ls + P - 32 - (int64_t)a.cbase + 32 * kHaloL;
            uint32_t l[1], h[1], o[1];
            window_planes<1>(rd, bit, strand, l, h, o);
            const uint32_t lc = l[0] & care, hc = h[0] & care, oc = o[0] & care;
            const TileSink sink{HitSink{a.hkey, a.hmm, a.count, a.cap, t0 + ls, strand, 0u}, stage_ref(s_stage),
                                stage_low(strand, (int32_t)ls)};
            for (uint32_t p = 0; p < sd.parts; p++) {
                const uint32_t pm = sd.mask[p];
                if (oc & pm) continue;  // a genome N at a comparable position: the range cannot match exactly
                const uint32_t a0 = sd.first[p] & 0xFF, len = sd.first[p] >> 8;
                const uint32_t code =
                    __builtin_amdgcn_ubfe(l[0], a0, len) | (__builtin_amdgcn_ubfe(h[0], a0, len) << len);
                const uint32_t b = sd.obase[p] + code;
                const uint32_t end = s_offs[b + 1];
                for (uint32_t j = s_offs[b]; j < end; j++) {
                    const uint32_t r = s_ids[j];
                    const uint32_t x = __builtin_amdgcn_bitop3_b32(hc, s_rec[3 * r + 1], oc, kXorOr);
                    const uint32_t y = __builtin_amdgcn_bitop3_b32(lc, s_rec[3 * r], x, kXorOr);
                    const int32_t sc = (int32_t)(s_rec[3 * r + 2] + (uint32_t)__popc(y));
                    if (sc < 0) {
                        bool first = true;  // emitted from the first range that matches exactly
                        for (uint32_t q = 0; q < p; q++) first = first && (y & sd.mask[q]) != 0;
                        if (first)
                            emit_hit(sink, a.table[kClassHdr + 4 * r + 3],
                                     (uint32_t)(sc - (int32_t)s_rec[3 * r + 2]));
                    }
                }
            }
        });
}

template <int NW, bool GEN>
static void fused_variant(int variant, dim3 grid, const FusedArgs &a, hipStream_t s) {
    note_kernel("fused_kernel", NW, GEN, variant == 1 ? 8 : 32);
    if (variant == 1) hipLaunchKernelGGL((fused_kernel<NW, GEN, 8>), grid, dim3(kThreads), 0, s, a);
    else hipLaunchKernelGGL((fused_kernel<NW, GEN, 32>), grid, dim3(kThreads), 0, s, a);
}

hipError_t launch_fused(const DevGenome &g, const PatternDesc &pd, int nw, bool gen, uint32_t cbase,
                        const uint32_t *table, uint32_t nunits, uint32_t gbase, uint64_t *hkey, uint16_t *hmm,
                        unsigned long long *count, unsigned long long *cand_count, uint64_t cap, hipStream_t s,
                        int variant, uint32_t tile_stride, uint32_t max_tiles) {
    const uint64_t limit = g.total >= pd.P ? g.total - pd.P + 1 : 0;
    if (limit == 0 || nunits == 0) return hipSuccess;
    if (tile_stride == 0) return hipErrorInvalidValue;
    FusedArgs a{};
    a.g = g;
    a.pd = pd;
    a.cbase = cbase;
    a.nunits = nunits;
    a.gbase = gbase;
    a.tile_stride = tile_stride;
    a.table = table;
    a.limit = limit;
    a.hkey = hkey;
    a.hmm = hmm;
    a.count = count;
    a.cand_count = cand_count;
    a.cap = cap;
    const dim3 grid = tile_grid(limit, tile_stride, max_tiles);
    if (nw == 1) gen ? fused_variant<1, true>(variant, grid, a, s) : fused_variant<1, false>(variant, grid, a, s);
    else gen ? fused_variant<2, true>(variant, grid, a, s) : fused_variant<2, false>(variant, grid, a, s);
    return hipGetLastError();
}

hipError_t launch_seed(const DevGenome &g, const PatternDesc &pd, const SeedDesc &sd, uint32_t lds_bytes,
                       uint32_t cbase, const uint32_t *table, const uint32_t *seed, uint64_t *hkey, uint16_t *hmm,
                       unsigned long long *count, unsigned long long *cand_count, uint64_t cap, hipStream_t s,
                       uint32_t tile_stride, uint32_t max_tiles) {
    const uint64_t limit = g.total >= pd.P ? g.total - pd.P + 1 : 0;
    if (limit == 0 || sd.nrec == 0) return hipSuccess;
    if (tile_stride == 0 || sd.parts == 0 || sd.parts > (uint32_t)kSeedMaxParts) return hipErrorInvalidValue;
    SeedArgs a{};
    a.g = g;
    a.pd = pd;
    a.sd = sd;
    a.cbase = cbase;
    a.tile_stride = tile_stride;
    a.table = table;
    a.seed = seed;
    a.limit = limit;
    a.hkey = hkey;
    a.hmm = hmm;
    a.count = count;
    a.cand_count = cand_count;
    a.cap = cap;
    const dim3 grid = tile_grid(limit, tile_stride, max_tiles);
    note_kernel("fused_seed_kernel");
    hipLaunchKernelGGL(fused_seed_kernel, grid, dim3(kThreads), (lds_bytes + 15) / 16 * 16, s, a);
    return hipGetLastError();
}

}  // namespace cof
