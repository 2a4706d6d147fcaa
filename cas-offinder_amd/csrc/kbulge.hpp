// Shared parts of bulge_kernel, bulge_index_kernel and bulge_slice_kernel.
#pragma once
#include "kcommon.hpp"

namespace cof {

// ------------------------------------------------------- device bulge expansion
//
// bulge_kernel scores guides instead of compare keys: for each candidate it
// takes the window at NS alignments (shift u: window position cb[u] at bit 0)
// and, per guide, the mismatch masks M[u] of the guide at each alignment.  The
// variants of the bulge family of shift u >= 1 (DNA or RNA bulge of one size,
// every split j) count mm(j) = popc(left & bits[0,j)) + popc(right &
// ~bits[0,j+gap)) with {left, right} = {M[u], M[0]}, so popc(M[u] & M[0]) - gap
// is a lower bound of every mm(j) of the family: one AND and one popcount
// filter the whole family.  Shift 0 alone is the no-bulge key (exact).  A
// family that passes is bounded again on the two halves of its split range,
// and only then are its splits counted one by one and emitted through the
// family's variant entries (key index, threshold as bias; a key shared by
// several variants is emitted by one of them).

struct BulgeArgs {
    DevGenome g;
    PatternDesc pd;
    int32_t cb[kMaxShifts];    // window position of mask bit 0 at each shift
    uint32_t gap[kMaxShifts];  // RNA-bulge length of the family of each shift
    uint32_t rev;              // reversed PAM: left part at shift 0, right part at shift u
    uint32_t pad;              // largest extra padding of a guide (0: every guide has the launch's)
    uint32_t nunits;           // care classes
    uint32_t tile_stride;
    const uint32_t *table;
    uint64_t limit;
    uint64_t *hkey;
    uint16_t *hmm;
    unsigned long long *count;
    unsigned long long *cand_count;
    uint64_t cap;
};

// bits [0, x) of word w of an NW-word mask (x may be negative or past the end)
__device__ __forceinline__ uint32_t low_bits(int32_t x, int w) {
    const int32_t y = x - 32 * w;
    return y <= 0 ? 0u : y >= 32 ? ~0u : (1u << y) - 1u;
}

template <int NW>
__device__ __forceinline__ uint32_t popc_below(const uint32_t (&m)[NW], int32_t x) {
    uint32_t c = 0;
#pragma unroll
    for (int w = 0; w < NW; w++) c += (uint32_t)__popc(m[w] & low_bits(x, w));
    return c;
}

template <int NW>
__device__ __forceinline__ uint32_t popc_from(const uint32_t (&m)[NW], int32_t x) {
    uint32_t c = 0;
#pragma unroll
    for (int w = 0; w < NW; w++) c += (uint32_t)__popc(m[w] & ~low_bits(x, w));
    return c;
}

// The variants of one family for lanes whose filter passed.  fb = the
// family block [first split, split count, entry offset, segments, segment
// records]; bias = the family filter bias -(max threshold + gap + 1).  Each
// segment of splits [ja, jz] is bounded by popc(left below ja) +
// popc(left & right in [ja, jz + gap)) + popc(right from jz + gap) - gap;
// only the splits of a segment that passes are counted exactly.
template <int NW>
__device__ __forceinline__ void bulge_family(kptr_t tab, kptr_t fb, const uint32_t (&left)[NW],
                                             const uint32_t (&right)[NW], uint32_t gap, uint32_t bias,
                                             const TileSink &sink) {
    constexpr int SW = bulge_seg_words(NW);
    static_assert(kBulgeSegs == 4, "two halves of two segments");
    const int32_t j0 = (int32_t)fb[0];
    kptr_t ent = tab + fb[2];
    // bound of the splits of segments [s0, s1]: left below the first split,
    // left & right from it to the last split + gap (every position neither
    // below nor above), right from there on.  An unused segment record is
    // zero: the bound then counts fewer positions and stays a lower bound.
    auto bound = [&](int s0, int s1) {
        const kptr_t r0 = fb + 4 + s0 * SW, r1 = fb + 4 + s1 * SW;
        uint32_t b = bias;
#pragma unroll
        for (int w = 0; w < NW; w++) {
            const uint32_t below = r0[2 + w], above = r1[2 + 2 * NW + w];
            b += (uint32_t)__popc(left[w] & below);
            b += (uint32_t)__popc(__builtin_amdgcn_bitop3_b32(left[w], right[w], ~(below | above), kAnd3));
            b += (uint32_t)__popc(right[w] & above);
        }
        return (int32_t)b;
    };
    // each half (two segments, an unused one has no splits) is bounded
    // first, its segments only when the half passes
#pragma unroll
    for (int hf = 0; hf < 2; hf++) {
        if (bound(2 * hf, 2 * hf + 1) >= 0) continue;
#pragma unroll
        for (int sg = 2 * hf; sg < 2 * hf + 2; sg++) {
            const kptr_t sr = fb + 4 + sg * SW;
            const uint32_t a = sr[0], cnt = sr[1];
            if (cnt == 0 || bound(sg, sg) >= 0) continue;
            for (uint32_t k = a; k < a + cnt; k++) {
                const int32_t j = j0 + (int32_t)k;
                const uint32_t mm = popc_below<NW>(left, j) + popc_from<NW>(right, j + (int32_t)gap);
                if ((int32_t)(ent[2 * k] + mm) < 0) emit_hit(sink, ent[2 * k + 1], mm);
            }
        }
    }
}

// bit x of an NW-word mask (0 for x outside it)
template <int NW>
__device__ __forceinline__ uint32_t mask_bit(const uint32_t (&m)[NW], int32_t x) {
    uint32_t b = 0;
#pragma unroll
    for (int w = 0; w < NW; w++)
        if ((uint32_t)(x - 32 * w) < 32u) b = (m[w] >> (x - 32 * w)) & 1u;
    return b;
}

// The variants of one family, every split counted in turn: for lanes that
// each hold their own (candidate, guide) pair (bulge_slice_kernel's queue),
// where the segment bounds of bulge_family diverge across the lanes and the
// wave runs every lane's path; here the lanes share one loop of at most the
// family's split count.  The count of split j + 1 is the count of split j
// plus left's bit j minus right's bit j + gap; a split's entry is read only
// when its count is within the family's largest threshold (bias + gap + mm
// < 0).
template <int NW>
__device__ __forceinline__ void bulge_family_flat(kptr_t tab, kptr_t fb, const uint32_t (&left)[NW],
                                                  const uint32_t (&right)[NW], uint32_t gap, uint32_t bias,
                                                  const TileSink &sink) {
    const int32_t j0 = (int32_t)fb[0];
    const uint32_t n = fb[1];
    const kptr_t ent = tab + fb[2];
    uint32_t mm = popc_below<NW>(left, j0) + popc_from<NW>(right, j0 + (int32_t)gap);
    const uint32_t bg = bias + gap;
    for (uint32_t k = 0; k < n; k++) {
        const int32_t j = j0 + (int32_t)k;
        if ((int32_t)(bg + mm) < 0 && (int32_t)(ent[2 * k] + mm) < 0) emit_hit(sink, ent[2 * k + 1], mm);
        mm += mask_bit<NW>(left, j) - mask_bit<NW>(right, j + (int32_t)gap);
    }
}

// Separator bitmap of a staged tile, for tables with guides of extra
// padding: bit b of word w stands for genome position t0 - 32 * kHaloL +
// 32 * w + b (the positions of the staged planes) and is set where the
// genome holds a separator (a poison position) or ends (positions below 0
// and from g.total on).  Thread t writes words t, t + kThreads, ... (read
// after the tile pipeline's first barrier).
constexpr uint32_t kSepBytes = kLdsWords * 4;
__device__ __forceinline__ void stage_separators(const DevGenome &g, uint64_t tile, uint32_t *s_sep) {
    const int64_t base = (int64_t)tile * kTilePos - 32 * kHaloL;
    const uint64_t lo = base > 0 ? (uint64_t)base : 0, hi = (uint64_t)(base + 32 * kLdsWords);
    const int t = (int)threadIdx.x;
    uint32_t m[kStageIters] = {};
    uint64_t i = g.tile_pidx[tile ? tile - 1 : 0];  // first separator at or after the previous tile's start
    for (;; i++) {
        const uint64_t x = g.poison[i];  // poison[npoison] is the sentinel
        if (x >= hi) break;
        if (x < lo) continue;
        const uint32_t b = (uint32_t)((int64_t)x - base), w = b >> 5;
#pragma unroll
        for (int k = 0; k < kStageIters; k++)
            if (w == (uint32_t)(t + k * kThreads)) m[k] |= 1u << (b & 31);
    }
#pragma unroll
    for (int k = 0; k < kStageIters; k++) {
        const int w = t + k * kThreads;
        const int64_t ws = base + 32 * w;  // position of bit 0
        if (ws < 0) m[k] |= ws <= -32 ? ~0u : (1u << -ws) - 1;
        const int64_t te = (int64_t)g.total - ws;  // first bit at or past the genome end
        if (te < 32) m[k] |= te <= 0 ? ~0u : ~((1u << te) - 1);
        if (w < kLdsWords) s_sep[w] = m[k];
    }
}

// A bulge table may hold queries of one pattern with different DNA-bulge
// sizes D: the launch finds the windows of the smallest padding (the
// pattern with the fewest padding N), and a guide of a larger D carries the
// extra padding d in its variant block word (vw >> kVbPadShift; the PAD
// instances run those tables, with the tile's separator bitmap s_sep).
// Where the padding lies before the window in the genome (forward PAM on the
// + strand, reversed PAM on the - strand) the guide's window starts d
// positions earlier; on the other side it ends d positions later.  The
// pair has no record when those d positions leave the genome or hold a
// separator (cas-offinder.cl:64-67 for the longer pattern).
//
// pad_room: free positions (at most kMaxShifts) beside candidate ce's window
// on the padding side, from one funnel of the bitmap: the pair of a guide of
// extra padding d has a record only if d <= room.
template <bool PAD>
__device__ __forceinline__ uint32_t pad_room(const BulgeArgs &a, uint32_t ce, const uint32_t *s_sep) {
    if (!PAD) return 0;
    const uint32_t off = ce >> 1, strand = ce & 1;
    const bool before = (strand == 0) != (a.rev != 0);
    // bit 0: the first of the kMaxShifts positions before / after the window
    // (bitmap bits [32 * kHaloL - kMaxShifts, 32 * (kLdsWords - 1)) for every candidate)
    const uint32_t b = (before ? off - kMaxShifts : off + a.pd.P) + 32 * kHaloL;
    const uint32_t x = funnel(s_sep[(b >> 5) + 1], s_sep[b >> 5], b & 31) & ((1u << kMaxShifts) - 1);
    // before: positions nearest the window are the high bits; after: the low ones
    return before ? (uint32_t)__clz(x) - (32 - kMaxShifts) : (uint32_t)__builtin_ctz(x | 1u << kMaxShifts);
}

// The sink of the pair (candidate entry ce, guide whose variant block word
// is vw = block offset | extra padding << kVbPadShift): the record starts at
// the guide's own window start.
template <bool PAD>
__device__ __forceinline__ TileSink bulge_sink(const BulgeArgs &a, uint64_t t0, uint32_t ce, kptr_t tab, uint32_t vw,
                                               StageRef st) {
    const uint32_t off = ce >> 1, strand = ce & 1;
    const uint32_t gbase = tab[vw & kVbOffsetMask];
    const uint32_t sh = PAD && (strand == 0) != (a.rev != 0) ? vw >> kVbPadShift : 0u;
    return TileSink{HitSink{a.hkey, a.hmm, a.count, a.cap, t0 + off - sh, strand, gbase}, st,
                    stage_low(strand, (int32_t)off - (int32_t)sh)};
}

// Window planes of candidate entry e (tile offset << 1 | strand) at the NS
// alignments of the bulge table.
template <int NW, int NS, typename Rd>
__device__ __forceinline__ void bulge_windows(const BulgeArgs &a, Rd rd, uint32_t e, uint32_t (&l)[NS][NW],
                                              uint32_t (&h)[NS][NW], uint32_t (&o)[NS][NW]) {
    const uint32_t ls = e >> 1, strand = e & 1;
#pragma unroll
    for (int u = 0; u < NS; u++) {
        const int64_t bit = strand == 0 ? (int64_t)ls + a.cb[u] + 32 * kHaloL
                                        : (int64_t)ls + a.pd.P - 32 * NW - (int64_t)a.cb[u] + 32 * kHaloL;
        window_planes<NW>(rd, bit, strand, l[u], h[u], o[u]);
    }
}

template <int NW>
__device__ __forceinline__ void class_planes(const uint32_t (&l)[NW], const uint32_t (&h)[NW], const uint32_t (&o)[NW],
                                             uint32_t (&cls)[5][NW]) {
#pragma unroll
    for (int w = 0; w < NW; w++) {
        cls[kClsT][w] = ~(l[w] | h[w] | o[w]);
        cls[kClsC][w] = l[w] & ~h[w];
        cls[kClsA][w] = h[w] & ~l[w];
        cls[kClsG][w] = l[w] & h[w];
        cls[kClsO][w] = o[w];
    }
}

// Mismatch mask of a general (IUPAC) guide record against the class planes.
template <int NW>
__device__ __forceinline__ uint32_t general_mask(kptr_t kp, const uint32_t (&cls)[5][NW], int w) {
    uint32_t x = cls[kClsO][w] & kp[kClsO * NW + w];
    x = __builtin_amdgcn_bitop3_b32(cls[kClsG][w], kp[kClsG * NW + w], x, kAndOr);
    x = __builtin_amdgcn_bitop3_b32(cls[kClsA][w], kp[kClsA * NW + w], x, kAndOr);
    x = __builtin_amdgcn_bitop3_b32(cls[kClsC][w], kp[kClsC * NW + w], x, kAndOr);
    return __builtin_amdgcn_bitop3_b32(cls[kClsT][w], kp[kClsT * NW + w], x, kAndOr);
}

}  // namespace cof
