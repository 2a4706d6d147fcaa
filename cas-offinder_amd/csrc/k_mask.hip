// Batched key tables behind the bit-set seed index: fused_mask_kernel.
#include "kcommon.hpp"

namespace cof {

// fused_mask_kernel: a single-class concrete key table of at most 32 * NK
// records with a bit-set seed index (MaskDesc).  Per candidate, each range's
// 2-bit codes select a row of NK words (bit r: record r's bases there equal
// the window's); the rows of all ranges are combined into the records that
// match at least MINEX ranges exactly (OR for 1; for 2, bits seen twice; for
// 3, a bit-sliced count per word against the word's need), and only those
// records are counted, each once.  A window within threshold t of a key
// matches at least parts - t ranges exactly, at least what its word asks
// for, so no hit is missed.  Compared with fused_seed_kernel there are no per-range lists to
// walk: the row loads are independent and the per-lane loop runs over the
// few surviving records.

struct MaskArgs {
    DevGenome g;
    PatternDesc pd;
    MaskDesc md;
    uint32_t cbase, tile_stride;
    uint32_t ntile;       // tiles of the launch (the workgroups hold G of them)
    const uint32_t *idx;  // rows then records (md.nwords words)
    uint64_t limit;
    uint64_t *hkey;
    uint16_t *hmm;
    unsigned long long *count;
    unsigned long long *cand_count;
    uint64_t cap;
};



// G tiles per workgroup (groups of kThreads threads) share one copy of the
// index in LDS; with G > 1 the candidate list holds kListCap entries.
// Register budget: six waves per SIMD, five for 12-word rows of up to 10
// ranges (they fit 96 VGPRs without spilling), or three for wider rows,
// whose LDS index allows no more waves anyway; with per-word
// counts their rows are read and counted for all words first, so that the
// reads of every chunk issue before the survivor loops.
template <int NK, int MINEX, int NP, int G>
__global__ __launch_bounds__(kThreads * G, NK >= kMaskWideWords ? (NK <= 12 && NP <= 10 ? 5 : 3) : 6) void fused_mask_kernel(MaskArgs a) {
    constexpr int LC = G == 1 ? kSeedListCap : kListCap;
    __shared__ uint32_t s_lo[G][kLdsWords], s_hi[G][kLdsWords], s_ot[G][kLdsWords];
    __shared__ uint16_t s_list[G][LC];
    __shared__ uint32_t s_wsum[G][4];
    __shared__ uint32_t s_nmax[G];
    __shared__ HitStageBuf<kSeedStageCap> s_stage[G];
    extern __shared__ __align__(16) uint32_t s_idx[];  // rows, then records at md.recw
    const MaskDesc &md = a.md;
    for (uint32_t i = threadIdx.x; i < md.nwords / 4; i += kThreads * G)
        reinterpret_cast<uint4 *>(s_idx)[i] = reinterpret_cast<const uint4 *>(a.idx)[i];
    // visible after tile_candidates' first barrier
    const uint32_t grp = G == 1 ? 0u : threadIdx.x / kThreads;
    uint32_t *g_lo = s_lo[grp], *g_hi = s_hi[grp], *g_ot = s_ot[grp];
    HitStageBuf<kSeedStageCap> &g_stage = s_stage[grp];
    const uint32_t *s_rows = s_idx;
    const uint4 *s_rec = reinterpret_cast<const uint4 *>(s_idx + md.recw);
    const uint32_t care = md.care;
    const uint32_t plane = mask_plane(md.zrow);  // words between chunk planes (rows of 4 or more words)
    const uint32_t fwd_bit = a.cbase + 32 * kHaloL, rev_bit = a.pd.P - 32 - a.cbase + 32 * kHaloL;
    tile_candidates<1, false, LC, G>(
        a.g, a.pd, a.limit, a.tile_stride, a.cand_count, g_lo, g_hi, g_ot, s_list[grp], s_wsum[grp],
        stage_ref(g_stage), a.hkey, a.hmm, a.count, a.cap,
        [&](uint64_t t0, auto, const uint32_t (&e)[1], const bool (&)[1]) {
            const uint32_t ls = e[0] >> 1, strand = e[0] & 1;
            uint32_t l, h, o;
            const uint32_t rsel = (uint32_t)((int32_t)(e[0] << 31) >> 31);  // all ones on reverse
            window_lds(g_lo, g_hi, g_ot, (rsel & (rev_bit - fwd_bit)) + ls + fwd_bit, strand, l, h, o);
            const uint32_t lc = l & care, hc = h & care, oc = o & care;
            // rows of every range first (independent LDS reads), then combine;
            // parts past md.parts have span 0 and address the zero row.  A
            // genome N reads as base T (lo = hi = 0), so a range holding one
            // may select keys it does not match: extra survivors that the
            // exact count below rejects.  No hit is lost: the ranges that a
            // hit's key matches exactly hold no N.
            // row pointers: words 0-3 of the row (chunk c of a row of 4 or
            // more words at c * plane further, chunk-major layout)
            constexpr uint32_t RS = NK < 4 ? NK : 4;
            const uint32_t *row[NP];
#pragma unroll
            for (int p = 0; p < NP; p++) {
                const uint32_t f = md.first[p];  // bit | span << 8 | first row << 16
                const uint32_t a0 = f & 0xFF, span = (f >> 8) & 0xFF;
                const uint32_t code =
                    __builtin_amdgcn_ubfe(l, a0, span) | (__builtin_amdgcn_ubfe(h, a0, span) << span);
                row[p] = s_rows + (f >> 16) * RS + code * RS;  // the range's first row is uniform
            }
            auto count = [&](uint32_t r) {  // exact count of surviving record r
                const uint4 k = s_rec[r];
                const uint32_t y = __builtin_amdgcn_bitop3_b32(hc, k.y, oc, kXorOr);
                const uint32_t mm = (uint32_t)__popc(__builtin_amdgcn_bitop3_b32(lc, k.x, y, kXorOr));
                if ((int32_t)(k.z + mm) < 0) {  // the sink is built only on a hit
                    uint32_t ev = e[0];
                    asm volatile("" : "+v"(ev));  // keeps the sink's arithmetic on the (rare) hit path
                    emit_hit(TileSink{HitSink{a.hkey, a.hmm, a.count, a.cap, t0 + (ev >> 1), ev & 1, 0u},
                                      stage_ref(g_stage), stage_low(ev & 1, (int32_t)(ev >> 1))},
                             k.w, mm);
                }
            };
            // rows in chunks of CH words (bounded registers for wide rows):
            // the chunk's words of every range first (independent LDS
            // reads), then combined, then its survivors counted two words at
            // a time (fewer loops whose trip count is the wave's largest)
            constexpr int CH = NK < 4 ? NK : 4;
            if constexpr (NK >= kMaskWideWords && MINEX == 3) {
                uint32_t sel[NK];
#pragma unroll
                for (int c = 0; c < NK / 4; c++) {
                    if constexpr (MINEX == 3) {
                        // the chunk's four needs (4 bits each); all 15: no records
                        uint32_t nc = (md.need[c / 2] >> (16 * (c % 2))) & 0xFFFFu;
                        asm volatile("" : "+s"(nc));
                        if (nc == 0xFFFFu) {  // uniform: a branch per chunk keeps one
                            // chunk of rows live at a time (the reads of every chunk
                            // ahead of the counts spill)
#pragma unroll
                            for (int w = 0; w < 4; w++) sel[4 * c + w] = 0;
                            continue;
                        }
                    }
                    uint32_t m[NP][4];
#pragma unroll
                    for (int p = 0; p < NP; p++) load_words<4>(row[p] + c * plane, m[p]);
                    if constexpr (MINEX == 3) {
#pragma unroll
                        for (int w = 0; w < 4; w++) {
                            const uint32_t cw = 4 * c + w;
                            sel[cw] = at_least<NP, 4>(m, w, (md.need[cw / 8] >> (4 * (cw % 8))) & 15u);
                        }
                    } else {
                        uint32_t any[4], two[4];
                        combine_rows<NP, 4, MINEX == 2>(m, any, two);
#pragma unroll
                        for (int w = 0; w < 4; w++) sel[4 * c + w] = MINEX == 2 ? two[w] : any[w];
                    }
                }
#pragma unroll
                for (int w = 0; w < NK; w += 2)
                    for (uint64_t x = (uint64_t)sel[w] | (uint64_t)sel[w + 1] << 32; x; x &= x - 1)
                        count(32 * w + (uint32_t)__builtin_ctzll(x));
                return;
            }
#pragma unroll 1
            for (int c = 0; c < NK / CH; c++) {
                uint32_t m[NP][CH];
#pragma unroll
                for (int p = 0; p < NP; p++) load_words<CH>(row[p] + c * plane, m[p]);
                uint32_t any[CH], two[CH], sel[CH];
                if constexpr (MINEX == 3) {  // per-word counts
#pragma unroll
                    for (int w = 0; w < CH; w++) {
                        const uint32_t cw = CH * c + w;
                        sel[w] = at_least<NP, CH>(m, w, (md.need[cw / 8] >> (4 * (cw % 8))) & 15u);
                    }
                } else {
                    combine_rows<NP, CH, MINEX == 2>(m, any, two);
#pragma unroll
                    for (int w = 0; w < CH; w++) sel[w] = MINEX == 2 ? two[w] : any[w];
                }
                if constexpr (CH == 1) {
                    for (uint32_t x = sel[0]; x; x &= x - 1) count(__builtin_ctz(x));
                } else {
#pragma unroll
                    for (int w = 0; w < CH; w += 2) {
                        for (uint64_t x = (uint64_t)sel[w] | (uint64_t)sel[w + 1] << 32; x; x &= x - 1)
                            count(32 * (CH * c + w) + (uint32_t)__builtin_ctzll(x));
                    }
                }
            }
        },
        a.ntile, s_nmax);
}

template <int NK, int MINEX, int NP, int G>
static hipError_t mask_launch(const MaskArgs &a, uint32_t lds, hipStream_t s) {
    const dim3 grid((a.ntile + G - 1) / G);
    note_kernel("fused_mask_kernel", NK, MINEX, NP, G);
    hipLaunchKernelGGL((fused_mask_kernel<NK, MINEX, NP, G>), grid, dim3(kThreads * G), lds, s, a);
    return hipGetLastError();
}

// Dynamic LDS limit of the two- and three-group mask kernels (their static
// LDS plus an index of up to kMaskMaxIndexBytes exceeds 64 KiB).
template <int NK, int MINEX>
static hipError_t mask_attr() {
    hipError_t e = hipSuccess;
    const void *k[3] = {(const void *)fused_mask_kernel<NK, MINEX, 4, 3>, (const void *)fused_mask_kernel<NK, MINEX, 6, 3>,
                        (const void *)fused_mask_kernel<NK, MINEX, 8, 3>};
    for (const void *f : k)
        if (e == hipSuccess) e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMaskMaxIndexBytes);
    if constexpr (NK >= kMaskWideWords) {
        const void *k2[3] = {(const void *)fused_mask_kernel<NK, MINEX, 4, 2>,
                             (const void *)fused_mask_kernel<NK, MINEX, 6, 2>,
                             (const void *)fused_mask_kernel<NK, MINEX, 8, 2>};
        for (const void *f : k2)
            if (e == hipSuccess)
                e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMaskMaxIndexBytes);
    }
    if constexpr (MINEX == 3 && NK >= kMaskWideWords) {  // 10 and 12 rows
        const void *k3[2] = {(const void *)fused_mask_kernel<NK, MINEX, 10, 3>,
                             (const void *)fused_mask_kernel<NK, MINEX, 12, 3>};
        for (const void *f : k3)
            if (e == hipSuccess)
                e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMaskMaxIndexBytes);
        if constexpr (NK >= kMaskWideWords) {
            const void *k4[2] = {(const void *)fused_mask_kernel<NK, MINEX, 10, 2>,
                                 (const void *)fused_mask_kernel<NK, MINEX, 12, 2>};
            for (const void *f : k4)
                if (e == hipSuccess)
                    e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMaskMaxIndexBytes);
        }
    }
    return e;
}

hipError_t prepare_kernels() {
    hipError_t e = hipSuccess;
    hipError_t (*fs[])() = {mask_attr<1, 1>,  mask_attr<1, 2>,  mask_attr<1, 3>,  mask_attr<2, 1>,
                            mask_attr<2, 2>,  mask_attr<2, 3>,  mask_attr<4, 1>,  mask_attr<4, 2>,
                            mask_attr<4, 3>,  mask_attr<8, 1>,  mask_attr<8, 2>,  mask_attr<8, 3>,
                            mask_attr<12, 2>, mask_attr<12, 3>, mask_attr<16, 2>, mask_attr<16, 3>,
                            mask_attr<20, 2>, mask_attr<20, 3>, mask_attr<24, 2>, mask_attr<24, 3>};
    for (auto f : fs)
        if (e == hipSuccess) e = f();
    return e;
}

// Static LDS of fused_mask_kernel per group: tile planes, candidate list,
// scan words, stage (and the group's s_nmax word).
constexpr uint32_t mask_group_lds(int lc) {
    return 3 * kLdsWords * 4 + lc * 2 + 16 + (uint32_t)sizeof(HitStageBuf<kSeedStageCap>) + 4;
}

// Groups per workgroup: the count (1, 3, or 2 for wide rows) that keeps the
// most waves resident per CU under the LDS limit, the larger on a tie (fewer
// index copies, one-pass candidate lists).
static int mask_groups(uint32_t idx_bytes, bool wide) {
    constexpr uint32_t kCuLds = 160 * 1024;
    const uint32_t max_waves = wide ? 16 : 24;  // the registers of wide-row kernels allow four waves per SIMD
    auto waves = [&](uint32_t g) {
        const uint32_t lds = g * mask_group_lds(g == 1 ? kSeedListCap : kListCap) + idx_bytes;
        return std::min(max_waves, kCuLds / lds * 4 * g);
    };
    int best = 1;
    for (uint32_t g : {3u, 2u})
        if ((wide || g == 3) && waves(g) >= waves((uint32_t)best)) best = (int)g;
    return best;
}

template <int NK, int MINEX, int NP>
static hipError_t mask_groups_launch(const MaskArgs &a, uint32_t lds, hipStream_t s) {
    constexpr bool wide = NK >= kMaskWideWords;
    switch (mask_groups(lds, wide)) {
        case 3: return mask_launch<NK, MINEX, NP, 3>(a, lds, s);
        case 2:
            if constexpr (wide) return mask_launch<NK, MINEX, NP, 2>(a, lds, s);
            return hipErrorInvalidValue;
        default: return mask_launch<NK, MINEX, NP, 1>(a, lds, s);
    }
}

template <int NK, int NP>
static hipError_t mask_minex(const MaskArgs &a, uint32_t lds, hipStream_t s) {
    if constexpr (NP > kSeedMaxParts) {  // 10 or 12 rows: per-word counts only
        if (a.md.minex == 3) return mask_groups_launch<NK, 3, NP>(a, lds, s);
        return hipErrorInvalidValue;
    } else {
        if (a.md.minex == 2) return mask_groups_launch<NK, 2, NP>(a, lds, s);
        if (a.md.minex == 3) return mask_groups_launch<NK, 3, NP>(a, lds, s);
        if constexpr (NK <= 8) return mask_groups_launch<NK, 1, NP>(a, lds, s);  // wider tables: minex 2, 3 only
        return hipErrorInvalidValue;
    }
}

template <int NK>
static hipError_t mask_parts(const MaskArgs &a, uint32_t lds, hipStream_t s) {
    if (a.md.parts <= 4) return mask_minex<NK, 4>(a, lds, s);
    if (a.md.parts <= 6) return mask_minex<NK, 6>(a, lds, s);
    if (a.md.parts <= 8) return mask_minex<NK, 8>(a, lds, s);
    if constexpr (NK >= kMaskWideWords) {  // wide rows only (narrower ones would spill registers)
        if (a.md.parts <= 10) return mask_minex<NK, 10>(a, lds, s);
        return mask_minex<NK, 12>(a, lds, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_mask(const DevGenome &g, const PatternDesc &pd, const MaskDesc &md, uint32_t cbase,
                       const uint32_t *idx, uint64_t *hkey, uint16_t *hmm, unsigned long long *count,
                       unsigned long long *cand_count, uint64_t cap, hipStream_t s, uint32_t tile_stride,
                       uint32_t max_tiles) {
    const uint64_t limit = g.total >= pd.P ? g.total - pd.P + 1 : 0;
    if (limit == 0 || md.nrec == 0) return hipSuccess;
    // more than kSeedMaxParts ranges: per-word counts of wide tables only
    const uint32_t max_parts =
        md.minex == 3 && md.nk >= (uint32_t)kMaskWideWords ? (uint32_t)kMaskMaxParts : (uint32_t)kSeedMaxParts;
    if (tile_stride == 0 || md.parts == 0 || md.parts > max_parts || md.minex < 1 || md.minex > 3 ||
        md.nrec > 32 * md.nk || md.nwords % 4 || md.recw % 4 || md.recw + 4 * md.nrec != md.nwords ||
        (uintptr_t)idx % 16)
        return hipErrorInvalidValue;
    for (uint32_t p = 0; p < (uint32_t)kMaskMaxParts; p++) {
        const uint32_t span = (md.first[p] >> 8) & 0xFF, rbase = md.first[p] >> 16;
        if (p >= md.parts ? md.first[p] != md.zrow << 16  // padding: span 0 at the zero row
                          : span == 0 || span > (uint32_t)kMaskMaxSpan || (md.first[p] & 0xFF) + span > 32 ||
                                rbase + (1u << (2 * span)) > md.zrow)
            return hipErrorInvalidValue;
    }
    if ((uint64_t)(md.zrow + 1) * md.nk > md.recw || md.nwords * 4 > kMaskMaxIndexBytes) return hipErrorInvalidValue;
    MaskArgs a{};
    a.g = g;
    a.pd = pd;
    a.md = md;
    a.cbase = cbase;
    a.tile_stride = tile_stride;
    a.idx = idx;
    a.limit = limit;
    a.hkey = hkey;
    a.hmm = hmm;
    a.count = count;
    a.cand_count = cand_count;
    a.cap = cap;
    a.ntile = tile_grid(limit, tile_stride, max_tiles).x;
    const uint32_t lds = md.nwords * 4;
    switch (md.nk) {
        case 1: return mask_parts<1>(a, lds, s);
        case 2: return mask_parts<2>(a, lds, s);
        case 4: return mask_parts<4>(a, lds, s);
        case 8: return mask_parts<8>(a, lds, s);
        case 12: return mask_parts<12>(a, lds, s);
        case 16: return mask_parts<16>(a, lds, s);
        case 20: return mask_parts<20>(a, lds, s);
        case 24: return mask_parts<24>(a, lds, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace cof
