// gfx950 kernels of the search path.
//
// Genome planes (lo, hi, ot) are 32-position words; a 256-thread workgroup
// owns one tile of kTileWords = 1024 words (32768 window starts, four
// consecutive words per thread) and stages the tile plus kHaloL / kHaloR
// words of halo in LDS.  Each thread tests the PAM of the 32 starts of each of
// its words on both strands with bitwise class masks (finder_masks), so one
// lane does 32 positions of cas-offinder.cl:26-63 at once.  Windows holding a
// ';' (poison list) or running past the genome end are cleared
// (cas-offinder.cl:64-67).
//
//   finder_count_kernel, finder_write_kernel
//                    plain path: (locus, flag) entries like cas-offinder.cl:68-70,
//                    at offsets from a scan of per-tile counts
//   comparer_kernel  plain path: one compare key per launch over the
//                    candidate list (cas-offinder.cl:73-155)
//   fused_kernel     batched path: finder + compaction into LDS + every
//                    compare key of a pattern group per candidate, no
//                    candidate list in HBM
//   bulge_kernel     batched path for queries with DNA/RNA bulges: the
//                    variants are expanded on the device from the guides
//
// Mismatch counting works on NW 32-bit words of the window planes taken at
// key position cbase.  A key whose comparable bases are all A/C/G/T is
// (klo, khi, care): mismatches = popcount(((lo^klo) | (hi^khi) | ot) & care);
// a key with IUPAC codes is five class masks (the classes that mismatch at
// each position): popcount(OR_c class_c & mask_c).  The threshold is folded in
// as a bias -(threshold+1) added by the popcount, so a score < 0 is a hit.
// The '-' strand reverse complements the window (bit reverse, hi ^= 1) and
// compares the forward key, which counts the same mismatches as the reverse
// complemented key against the forward text (cas-offinder.cl:123-154).
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>

#include "devtypes.hpp"

namespace cof {

namespace {

// Key tables are read-only for a launch: reading them through the constant
// address space lets the compiler use scalar loads (one per wavefront).
typedef const __attribute__((address_space(4))) uint32_t *kptr_t;

__device__ __forceinline__ uint32_t funnel(uint32_t hw, uint32_t lw, uint32_t sh) {
    return __builtin_amdgcn_alignbit(hw, lw, sh);  // bits [sh, sh+32) of hw:lw
}

// Class match of the planes for a 5-bit class set M (bit kClsT..kClsO: the
// classes that MATCH) as one v_bitop3 over (lo, hi, ot), truth-table index
// lo*4 + hi*2 + ot: T = 000, other = 001 (ot is set only where lo = hi = 0),
// A = 010, C = 100, G = 110; the other indices never occur.
template <uint32_t M>
constexpr uint32_t match_tt() {
    return ((M >> kClsT) & 1) << 0 | ((M >> kClsO) & 1) << 1 | ((M >> kClsA) & 1) << 2 |
           ((M >> kClsC) & 1) << 4 | ((M >> kClsG) & 1) << 6;
}

// acc[j] &= class match, for the class set M, of the planes shifted to
// pattern position k (bit b of word j <-> start b of word i0 + j).  The match
// is bitwise, so it is taken on the unshifted words (one v_bitop3 each) and
// only the one match plane is shifted.
template <uint32_t M, int NJ, int NL>
__device__ __forceinline__ void and_match(uint32_t (&acc)[NJ], const uint32_t (&L)[NL], const uint32_t (&H)[NL],
                                          const uint32_t (&O)[NL], uint32_t k) {
    uint32_t mt[NL];
#pragma unroll
    for (int i = 0; i < NL; i++) mt[i] = __builtin_amdgcn_bitop3_b32(L[i], H[i], O[i], match_tt<M>());
    const uint32_t sh = k & 31;
    const bool up = k >= 32;
#pragma unroll
    for (int j = 0; j < NJ; j++) acc[j] &= funnel(up ? mt[j + 2] : mt[j + 1], up ? mt[j + 1] : mt[j], sh);
}

// The class set of a pattern position is uniform, so a scalar switch picks
// the truth table.
template <int NJ, int NL>
__device__ __forceinline__ void and_match_any(uint32_t m, uint32_t (&acc)[NJ], const uint32_t (&L)[NL],
                                              const uint32_t (&H)[NL], const uint32_t (&O)[NL], uint32_t k) {
#define COF_MATCH_CASE(x) \
    case x: and_match<x>(acc, L, H, O, k); break;
    switch (m & 31u) {
        COF_MATCH_CASE(0) COF_MATCH_CASE(1) COF_MATCH_CASE(2) COF_MATCH_CASE(3)
        COF_MATCH_CASE(4) COF_MATCH_CASE(5) COF_MATCH_CASE(6) COF_MATCH_CASE(7)
        COF_MATCH_CASE(8) COF_MATCH_CASE(9) COF_MATCH_CASE(10) COF_MATCH_CASE(11)
        COF_MATCH_CASE(12) COF_MATCH_CASE(13) COF_MATCH_CASE(14) COF_MATCH_CASE(15)
        COF_MATCH_CASE(16) COF_MATCH_CASE(17) COF_MATCH_CASE(18) COF_MATCH_CASE(19)
        COF_MATCH_CASE(20) COF_MATCH_CASE(21) COF_MATCH_CASE(22) COF_MATCH_CASE(23)
        COF_MATCH_CASE(24) COF_MATCH_CASE(25) COF_MATCH_CASE(26) COF_MATCH_CASE(27)
        COF_MATCH_CASE(28) COF_MATCH_CASE(29) COF_MATCH_CASE(30) COF_MATCH_CASE(31)
    }
#undef COF_MATCH_CASE
}

// PAM test of the kWordsPerThread consecutive LDS words starting at i0 (the
// two words after them give the pattern reach): bit b of fwd[j] / rev[j] =
// start b of word i0 + j passes the forward / reverse-complement pattern.
__device__ __forceinline__ void finder_masks(const PatternDesc &pd, const uint32_t *s_lo, const uint32_t *s_hi,
                                             const uint32_t *s_ot, int i0, uint32_t (&fwd)[kWordsPerThread],
                                             uint32_t (&rev)[kWordsPerThread]) {
    constexpr int NJ = kWordsPerThread, NL = kWordsPerThread + 2;
    uint32_t L[NL], H[NL], O[NL];
#pragma unroll
    for (int i = 0; i < NL; i++) {
        L[i] = s_lo[i0 + i];
        H[i] = s_hi[i0 + i];
        O[i] = s_ot[i0 + i];
    }
#pragma unroll
    for (int j = 0; j < NJ; j++) fwd[j] = rev[j] = ~0u;
    // the empty asm marks the planes as changed on every pass, so that the
    // class matches of the 32 switch cases are not hoisted out of the loops
    // (all of them live at once would spill)
    auto opaque = [&]() {
#pragma unroll
        for (int i = 0; i < NL; i++) asm volatile("" : "+v"(L[i]), "+v"(H[i]), "+v"(O[i]));
    };
    for (uint32_t i = 0; i < pd.nf; i++) {
        opaque();
        and_match_any(pd.f[i] >> 8, fwd, L, H, O, pd.f[i] & 0xFF);
    }
    for (uint32_t i = 0; i < pd.nr; i++) {
        opaque();
        and_match_any(pd.r[i] >> 8, rev, L, H, O, pd.r[i] & 0xFF);
    }
}

// Bits of the starts [s0, s0+32) inside [lo, hi).
__device__ __forceinline__ uint32_t range_mask(uint64_t s0, uint64_t lo, uint64_t hi) {
    uint32_t m = ~0u;
    if (s0 + 32 > hi) m = s0 >= hi ? 0u : (uint32_t)((1ull << (hi - s0)) - 1);
    if (s0 < lo) m &= lo >= s0 + 32 ? 0u : ~(uint32_t)((1ull << (lo - s0)) - 1);
    return m;
}

// Clears, in the masks of the words starting at s0[j], the starts whose window
// [s, s+P) holds a poison position.  The per-tile table gives the first poison
// at or after the tile start t0, so a tile without one costs one scalar load.
// ppos = g.tile_ppos[tile], read by the caller before staging the tile so
// that the two memory latencies overlap.
template <int NJ>
__device__ __forceinline__ void apply_poison(const DevGenome &g, uint32_t P, uint64_t tile, uint64_t t1,
                                             uint64_t ppos, const uint64_t (&s0)[NJ], uint32_t (&keep)[NJ]) {
    const uint64_t end = t1 + P - 1;  // poisons at or past `end` touch no start of this tile
    if (ppos >= end) return;
    for (uint64_t i = g.tile_pidx[tile];; i++) {
        const uint64_t x = g.poison[i];  // poison[npoison] is the sentinel
        if (x >= end) break;
        const uint64_t a = x + 1 >= P ? x + 1 - P : 0;  // starts [a, x] see x
#pragma unroll
        for (int j = 0; j < NJ; j++) {
            if (x >= s0[j] && a < s0[j] + 32) {
                const uint32_t lb = a > s0[j] ? (uint32_t)(a - s0[j]) : 0u;
                const uint32_t hb = x < s0[j] + 31 ? (uint32_t)(x - s0[j]) : 31u;
                keep[j] &= ~(uint32_t)(((2ull << hb) - 1) & ~((1ull << lb) - 1));
            }
        }
    }
}

// The planes of one tile with its halo (kLdsWords words each), spread over
// the workgroup: thread t holds LDS words t, t + kThreads, ...
constexpr int kStageIters = (kLdsWords + kThreads - 1) / kThreads;
struct TileRegs {
    uint32_t lo[kStageIters], hi[kStageIters], ot[kStageIters];
};

__device__ __forceinline__ void load_tile(const DevGenome &g, int64_t w0, TileRegs &r, int t) {
    if (w0 >= kHaloL && (uint64_t)(w0 - kHaloL + kLdsWords) <= g.nwords) {
        // whole tile and halo inside the planes (every tile but the first
        // and the last): one base per plane, 32-bit offsets
        const uint32_t *lo = g.lo + (w0 - kHaloL), *hi = g.hi + (w0 - kHaloL), *ot = g.ot + (w0 - kHaloL);
#pragma unroll
        for (int k = 0; k < kStageIters; k++) {
            const int i = t + k * kThreads;
            const bool in = k + 1 < kStageIters || i < kLdsWords;
            r.lo[k] = in ? lo[i] : 0u;
            r.hi[k] = in ? hi[i] : 0u;
            r.ot[k] = in ? ot[i] : ~0u;
        }
        return;
    }
#pragma unroll
    for (int k = 0; k < kStageIters; k++) {
        const int i = t + k * kThreads;
        const int64_t w = w0 - kHaloL + i;
        const bool in = i < kLdsWords && w >= 0 && (uint64_t)w < g.nwords;
        r.lo[k] = in ? g.lo[w] : 0u;
        r.hi[k] = in ? g.hi[w] : 0u;
        r.ot[k] = in ? g.ot[w] : ~0u;
    }
}

__device__ __forceinline__ void store_tile(const TileRegs &r, uint32_t *s_lo, uint32_t *s_hi, uint32_t *s_ot,
                                           int t) {
#pragma unroll
    for (int k = 0; k < kStageIters; k++) {
        const int i = t + k * kThreads;
        if (i < kLdsWords) {
            s_lo[i] = r.lo[k];
            s_hi[i] = r.hi[k];
            s_ot[i] = r.ot[k];
        }
    }
}

// t: the thread's index among the kThreads threads that stage the tile.
__device__ __forceinline__ void stage_tile(const DevGenome &g, int64_t w0, uint32_t *s_lo, uint32_t *s_hi,
                                           uint32_t *s_ot, int t = (int)threadIdx.x) {
    TileRegs r;
    load_tile(g, w0, r, t);
    store_tile(r, s_lo, s_hi, s_ot, t);
}

// Inclusive prefix sum over the 64 lanes of a wave with DPP moves: shifts
// by 1, 2, 4 and 8 lanes within each row of 16 (lanes shifted in from
// outside the row read 0), then row_bcast:15 adds each row's last lane to
// the next row and row_bcast:31 adds lane 31 to rows 2 and 3.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    constexpr int kRowShr = 0x110, kRowBcast15 = 0x142, kRowBcast31 = 0x143;
    uint32_t x = v;
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, kRowShr + 1, 0xF, 0xF, false);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, kRowShr + 2, 0xF, 0xF, false);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, kRowShr + 4, 0xF, 0xF, false);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, kRowShr + 8, 0xF, 0xF, false);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, kRowBcast15, 0xA, 0xF, false);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, kRowBcast31, 0xC, 0xF, false);
    return x;
}

// exclusive scan of `v` over the workgroup (four wave64s); *total gets the sum
__device__ __forceinline__ uint32_t block_scan(uint32_t v, uint32_t *s_wsum, uint32_t *total,
                                               int t = (int)threadIdx.x) {
    const int lane = t & 63, wave = t >> 6;
    const uint32_t x = wave_incl_scan(v);
    if (lane == 63) s_wsum[wave] = x;
    __syncthreads();
    uint32_t off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < 4; w++) {
        const uint32_t s = s_wsum[w];
        off += w < wave ? s : 0u;
        tot += s;
    }
    *total = tot;
    return off + x - v;
}

// Where a workgroup appends hits: a 64-bit sort key
// (global key index << 40 | window start << 1 | strand) and the mismatch count.
struct HitSink {
    uint64_t *hkey;
    uint16_t *hmm;
    unsigned long long *count;
    uint64_t cap;
    uint64_t pos;
    uint32_t strand, gbase;
};

__device__ __forceinline__ void emit_hit(const HitSink &o, uint32_t key, uint32_t mm) {
    const unsigned long long slot = atomicAdd(o.count, 1ull);
    if (slot < o.cap) {
        o.hkey[slot] = ((uint64_t)(o.gbase + key) << kHitKeyShift) | (o.pos << 1) | o.strand;
        o.hmm[slot] = (uint16_t)mm;
    }
}

// Hits of a workgroup are staged in LDS and written out once per candidate
// pass with one global atomic, so the hit counter sees one atomic per tile
// pass instead of one per hit (same-address atomics serialize in one L2
// channel).  An entry packs global key index (bits 23..46), mismatches
// (16..22), strand (15) and the start's offset in the tile (0..14).
struct HitStage {  // header of a workgroup's stage
    uint32_t n;       // entries appended this pass; past the capacity they went to HBM directly
    uint32_t pad;
    unsigned long long base;
};
template <int CAP>
struct HitStageBuf {
    HitStage h;
    uint64_t ent[CAP];
};
struct StageRef {
    HitStage *h;
    uint64_t *ent;
    uint32_t cap;
};
template <int CAP>
__device__ __forceinline__ StageRef stage_ref(HitStageBuf<CAP> &b) {
    return StageRef{&b.h, b.ent, (uint32_t)CAP};
}
constexpr int kStageCap = 512;  // default stage entries
static_assert(kTilePos <= (1 << 15) && kMaxPattern < (1 << 7), "stage entry fields");

struct TileSink {
    HitSink g;      // direct path when the stage is full
    StageRef st;
    uint32_t low;   // strand << 15 | tile offset
};

__device__ __forceinline__ void emit_hit(const TileSink &o, uint32_t key, uint32_t mm) {
    const uint32_t i = atomicAdd(&o.st.h->n, 1u);
    if (i < o.st.cap)
        o.st.ent[i] = ((uint64_t)(o.g.gbase + key) << 23) | (mm << 16) | o.low;
    else
        emit_hit(o.g, key, mm);
}

// Writes the staged hits of the pass (all threads, after a barrier) and
// empties the stage.
// t: the thread's index among the kThreads threads that share the stage.
__device__ __forceinline__ void flush_stage(StageRef st, uint64_t t0, uint64_t *hkey, uint16_t *hmm,
                                            unsigned long long *count, uint64_t cap, int t = (int)threadIdx.x) {
    const uint32_t n = min(st.h->n, st.cap);
    if (t == 0) st.h->base = n ? atomicAdd(count, (unsigned long long)n) : 0ull;
    __syncthreads();
    const unsigned long long base = st.h->base;
    for (uint32_t i = (uint32_t)t; i < n; i += kThreads) {
        const uint64_t e = st.ent[i];
        const unsigned long long o = base + i;
        if (o < cap) {
            hkey[o] = ((e >> 23) << kHitKeyShift) | ((t0 + (e & 0x7FFF)) << 1) | ((e >> 15) & 1);
            hmm[o] = (uint16_t)((e >> 16) & 0x7F);
        }
    }
    __syncthreads();
    if (t == 0) st.h->n = 0;
}

// NW words of the window planes starting at plane bit `bit` (rd(i) reads word
// i of the three planes); for strand 1 the words are reverse complemented.
// Results have lo = hi = 0 wherever ot = 1.
template <int NW, typename Rd>
__device__ __forceinline__ void window_planes(Rd rd, int64_t bit, uint32_t strand, uint32_t (&l)[NW],
                                              uint32_t (&h)[NW], uint32_t (&o)[NW]) {
#pragma unroll
    for (int w = 0; w < NW; w++) {
        const int64_t b = bit + 32 * w;
        const int64_t wi = b >> 5;
        const uint32_t sh = (uint32_t)(b & 31);
        uint32_t a0, a1, b0, b1, c0, c1;
        rd(wi, a0, b0, c0);
        rd(wi + 1, a1, b1, c1);
        l[w] = funnel(a1, a0, sh);
        h[w] = funnel(b1, b0, sh);
        o[w] = funnel(c1, c0, sh);
    }
    if (strand) {
        uint32_t l2[NW], h2[NW], o2[NW];
#pragma unroll
        for (int w = 0; w < NW; w++) {
            l2[w] = __builtin_bitreverse32(l[NW - 1 - w]);
            o2[w] = __builtin_bitreverse32(o[NW - 1 - w]);
            h2[w] = ~__builtin_bitreverse32(h[NW - 1 - w]) & ~o2[w];  // complement: hi ^= 1 on bases
        }
#pragma unroll
        for (int w = 0; w < NW; w++) { l[w] = l2[w]; h[w] = h2[w]; o[w] = o2[w]; }
    }
}

// v_bitop3_b32 truth tables (bit index = a*4 + b*2 + c)
constexpr uint32_t kXorOr = 0xBE;  // (a ^ b) | c
constexpr uint32_t kAndOr = 0xEA;  // (a & b) | c
constexpr uint32_t kAnd3 = 0x80;   // a & b & c

// Concrete key table (every comparable key base is A/C/G/T): a run of care
// classes, each a header [care[NW], nkeys, 0...] of kClassHdr dwords followed
// by nkeys records [klo[NW], khi[NW], bias, key index] (klo/khi masked by
// care).  With the window masked once per class (lc = lo & care, ...), a key
// costs two v_bitop3 and one v_bcnt:
//   mismatches = popcount((lc ^ klo) | (hc ^ khi) | oc).
// General table (IUPAC codes): records [mask[class][NW] for T, C, A, G, other,
// bias, key index]: mismatches = popcount(OR_class plane_class & mask_class).
// bias = -(threshold + 1): popcount + bias < 0 is a hit.
template <int NW, bool GEN>
struct KeyFmt {
    static constexpr int stride = key_stride(NW, GEN);
    static constexpr int bias = key_bias_index(NW, GEN);
};

struct alignas(16) Rec4 {
    uint32_t x, y, z, w;
};
typedef const __attribute__((address_space(4))) Rec4 *kptr4_t;

// Concrete record read as aligned 16-byte rows (so eight records are two
// s_load_dwordx16): NW = 1 is {klo, khi, bias, index}; NW = 2 is
// {klo0, klo1, khi0, khi1}, {bias, index, 0, 0}.
template <int NW>
__device__ __forceinline__ int32_t score_concrete(kptr_t kw, const uint32_t (&lc)[NW], const uint32_t (&hc)[NW],
                                                  const uint32_t (&oc)[NW]) {
    const kptr4_t r = (kptr4_t)kw;
    if (NW == 1) {
        const uint32_t klo = r[0].x, khi = r[0].y, bias = r[0].z;
        const uint32_t x = __builtin_amdgcn_bitop3_b32(hc[0], khi, oc[0], kXorOr);  // (hc^khi) | oc
        return (int32_t)(bias + (uint32_t)__popc(__builtin_amdgcn_bitop3_b32(lc[0], klo, x, kXorOr)));
    } else {
        const uint32_t klo0 = r[0].x, klo1 = r[0].y, khi0 = r[0].z, khi1 = r[0].w, bias = r[1].x;
        const uint32_t x0 = __builtin_amdgcn_bitop3_b32(hc[0], khi0, oc[0], kXorOr);
        const uint32_t x1 = __builtin_amdgcn_bitop3_b32(hc[NW - 1], khi1, oc[NW - 1], kXorOr);
        const uint32_t acc = bias + (uint32_t)__popc(__builtin_amdgcn_bitop3_b32(lc[0], klo0, x0, kXorOr));
        return (int32_t)(acc + (uint32_t)__popc(__builtin_amdgcn_bitop3_b32(lc[NW - 1], klo1, x1, kXorOr)));
    }
}

template <int NW>
__device__ __forceinline__ int32_t score_general(kptr_t kw, const uint32_t (&cls)[5][NW]) {
    uint32_t acc = kw[5 * NW];
#pragma unroll
    for (int w = 0; w < NW; w++) {
        uint32_t x = cls[kClsO][w] & kw[kClsO * NW + w];
        x = __builtin_amdgcn_bitop3_b32(cls[kClsG][w], kw[kClsG * NW + w], x, kAndOr);
        x = __builtin_amdgcn_bitop3_b32(cls[kClsA][w], kw[kClsA * NW + w], x, kAndOr);
        x = __builtin_amdgcn_bitop3_b32(cls[kClsC][w], kw[kClsC * NW + w], x, kAndOr);
        x = __builtin_amdgcn_bitop3_b32(cls[kClsT][w], kw[kClsT * NW + w], x, kAndOr);
        acc += (uint32_t)__popc(x);
    }
    return (int32_t)acc;
}

// Scores n key records at kp against W windows at once (each scalar-loaded
// key is applied to every window: register blocking).  score(kw, w) is the
// biased score of window w; record stride S, bias at B, key index at B + 1.
// Scores of 32 keys are OR-ed into one sign test; only a block in which some
// lane has a hit is scored again key by key to emit.  Window w emits only if
// valid[w] (an invalid slot duplicates a valid window of the same lane).
template <int W, int S, int B, typename Score, typename Sink>
__device__ __forceinline__ void emit_block(kptr_t kp, uint32_t n, const Score &score, const Sink (&o)[W],
                                           const bool (&valid)[W]) {
    for (uint32_t k = 0; k < n; k++, kp += S) {
#pragma unroll
        for (int w = 0; w < W; w++) {
            const int32_t d = score(kp, w);
            if (d < 0 && valid[w]) emit_hit(o[w], kp[B + 1], (uint32_t)(d - (int32_t)kp[B]));
        }
    }
}


template <int W, int S, int B, int BLK, typename Score>
__device__ __forceinline__ bool block_hit(kptr_t kp, const Score &score) {
    int32_t acc = 0;
#pragma unroll
    for (int j = 0; j < BLK; j++) {
        int32_t d = score(kp + j * S, 0);
#pragma unroll
        for (int w = 1; w < W; w++) d |= score(kp + j * S, w);
        acc |= d;
    }
    return acc < 0;
}

// Tables of the fused kernel hold a multiple of 8 records per run (padded
// with records that never hit), so only the plain path's one-key tables reach
// the per-key tail.
// BLK keys share one sign test (32 or 8: larger blocks test less often,
// smaller ones rescore fewer keys when a block holds a hit).
template <int W, int S, int B, int BLK, typename Score, typename Sink>
__device__ __forceinline__ void score_run(kptr_t kp, uint32_t n, const Score &score, const Sink (&o)[W],
                                          const bool (&valid)[W]) {
    uint32_t k = 0;
    for (; k + BLK <= n; k += BLK, kp += BLK * S)
        if (block_hit<W, S, B, BLK>(kp, score)) emit_block<W, S, B>(kp, BLK, score, o, valid);
    for (; k + 8 <= n; k += 8, kp += 8 * S)
        if (block_hit<W, S, B, 8>(kp, score)) emit_block<W, S, B>(kp, 8, score, o, valid);
    if (k < n) emit_block<W, S, B>(kp, n - k, score, o, valid);
}

// Scores a whole key table (nunits = care classes for concrete tables, keys
// for general ones) against W windows (planes l, h, o of each).
template <int NW, bool GEN, int W, int BLK = 32, typename Sink>
__device__ __forceinline__ void score_table(kptr_t kp, uint32_t nunits, const uint32_t (&l)[W][NW],
                                            const uint32_t (&h)[W][NW], const uint32_t (&o)[W][NW],
                                            const Sink (&sink)[W], const bool (&valid)[W]) {
    constexpr int S = KeyFmt<NW, GEN>::stride, B = KeyFmt<NW, GEN>::bias;
    if (GEN) {
        uint32_t cls[W][5][NW];
#pragma unroll
        for (int w = 0; w < W; w++)
#pragma unroll
            for (int x = 0; x < NW; x++) {
                cls[w][kClsT][x] = ~(l[w][x] | h[w][x] | o[w][x]);
                cls[w][kClsC][x] = l[w][x] & ~h[w][x];
                cls[w][kClsA][x] = h[w][x] & ~l[w][x];
                cls[w][kClsG][x] = l[w][x] & h[w][x];
                cls[w][kClsO][x] = o[w][x];
            }
        score_run<W, S, B, BLK>(kp, nunits, [&](kptr_t kw, int w) { return score_general<NW>(kw, cls[w]); }, sink,
                                valid);
    } else {
        for (uint32_t c = 0; c < nunits; c++) {
            uint32_t lc[W][NW], hc[W][NW], oc[W][NW];
#pragma unroll
            for (int x = 0; x < NW; x++) {
                const uint32_t care = kp[x];
#pragma unroll
                for (int w = 0; w < W; w++) {
                    lc[w][x] = l[w][x] & care;
                    hc[w][x] = h[w][x] & care;
                    oc[w][x] = o[w][x] & care;
                }
            }
            const uint32_t n = kp[NW];
            kp += kClassHdr;
            score_run<W, S, B, BLK>(
                kp, n, [&](kptr_t kw, int w) { return score_concrete<NW>(kw, lc[w], hc[w], oc[w]); }, sink, valid);
            kp += n * S;
        }
    }
}

}  // namespace

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

constexpr int kListCap = 4096;  // LDS candidate entries per pass

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

// --------------------------------------------------------------- batched path

// Finder stage of the batched kernels for the tile of this workgroup: PAM
// masks, poison and range clearing, compaction of (start, strand) entries into
// an LDS list of LC entries per pass, then cands(t0, rd, entry, valid)
// per lane for chunks of 64 * W candidates (round-robin over the four waves).
// Entries are tile offset << 1 | strand (forward entries of a word first).
// Hits go through the workgroup's LDS stage st, flushed to (hkey, hmm,
// count, cap) after every pass.
// With ALL, cands runs on every lane of a wave whose chunk holds a candidate
// (lanes past the list end get valid = false), so it may use wave-wide
// ballots and cross-lane queues.
// The workgroup's tile is blockIdx.x * tile_stride (a stride above 1 samples
// tiles spread over the genome: autotuner trials).
// With G > 1 the workgroup holds G groups of kThreads threads, each with its
// own tile (tile index blockIdx.x * G + group, below ntile) and LDS arrays
// (the pointers passed are the group's); the groups run the same number of
// candidate passes, so that every barrier is reached by all of them.
// s_nmax: G words of LDS (G > 1 only).
template <int W, bool ALL = false, int LC = kListCap, int G = 1, typename Cands>
__device__ __forceinline__ void tile_candidates(const DevGenome &g, const PatternDesc &pd, uint64_t limit,
                                                uint32_t tile_stride, unsigned long long *cand_count,
                                                uint32_t *s_lo, uint32_t *s_hi, uint32_t *s_ot, uint16_t *s_list,
                                                uint32_t *s_wsum, StageRef st, uint64_t *hkey, uint16_t *hmm,
                                                unsigned long long *count, uint64_t cap, Cands &&cands,
                                                uint32_t ntile = 0, uint32_t *s_nmax = nullptr) {
    const int t = G == 1 ? (int)threadIdx.x : (int)(threadIdx.x & (kThreads - 1));
    const uint32_t grp = G == 1 ? 0u : threadIdx.x / kThreads;
    const uint32_t gi = G == 1 ? blockIdx.x : blockIdx.x * G + grp;
    const bool live = G == 1 || gi < ntile;
    const uint32_t tile = live ? gi * tile_stride : 0u;
    const int64_t w0 = (int64_t)tile * kTileWords;
    const uint64_t ppos = g.tile_ppos[tile];
    stage_tile(g, w0, s_lo, s_hi, s_ot, t);
    if (t == 0) st.h->n = 0;
    __syncthreads();
    const uint64_t t0 = (uint64_t)w0 * 32, t1 = t0 + kTilePos;
    uint32_t fwd[kWordsPerThread], rev[kWordsPerThread], keep[kWordsPerThread];
    uint64_t s0[kWordsPerThread];
    finder_masks(pd, s_lo, s_hi, s_ot, kWordsPerThread * t + kHaloL, fwd, rev);  // words 4t .. 4t+3
    const bool whole = t1 <= limit;  // every start of the tile is valid (all tiles but the last)
#pragma unroll
    for (int j = 0; j < kWordsPerThread; j++) {
        s0[j] = t0 + 32ull * (kWordsPerThread * t + j);
        keep[j] = whole ? ~0u : range_mask(s0[j], 0, limit);
    }
    apply_poison(g, pd.P, tile, t1, ppos, s0, keep);
    uint32_t mine = 0;
#pragma unroll
    for (int j = 0; j < kWordsPerThread; j++) {
        if (!live) keep[j] = 0;
        fwd[j] &= keep[j];
        rev[j] &= keep[j];
        mine += __popc(fwd[j]) + __popc(rev[j]);
    }
    uint32_t n;
    const uint32_t first = block_scan(mine, s_wsum, &n, t);
    if (t == 0 && n) atomicAdd(cand_count + (gi % kCountSlots) * kSlotStride, (unsigned long long)n);
    uint32_t nall = n;  // passes of the workgroup: the largest candidate count of its groups
    if (G > 1) {
        if (t == 0) s_nmax[grp] = n;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < G; k++) nall = max(nall, s_nmax[k]);
    }

    const int lane = t & 63, wave = t >> 6;
    auto rd = [=](int64_t wi, uint32_t &l, uint32_t &h, uint32_t &o) {
        l = s_lo[wi];
        h = s_hi[wi];
        o = s_ot[wi];
    };
    for (uint32_t pass = 0; pass < nall; pass += LC) {
        const uint32_t cnt = pass < n ? min(n - pass, (uint32_t)LC) : 0u;
        if (first < pass + cnt && first + mine > pass) {  // this thread has entries in the pass
            // entry of bit b of word wi: (wi * 32 + b) << 1 | strand; idx counts
            // from the pass start (it wraps below it: then idx >= cnt)
            uint32_t idx = first - pass;
            auto put = [&](bool all) {
#pragma unroll
                for (int j = 0; j < kWordsPerThread; j++) {
                    const uint32_t base = (uint32_t)(kWordsPerThread * t + j) << 6;
                    for (uint32_t m = fwd[j]; m; m &= m - 1, idx++)
                        if (all || idx < cnt) s_list[idx] = (uint16_t)(base + 2 * __builtin_ctz(m));
                    for (uint32_t m = rev[j]; m; m &= m - 1, idx++)
                        if (all || idx < cnt) s_list[idx] = (uint16_t)(base + 2 * __builtin_ctz(m) + 1);
                }
            };
            if (n <= (uint32_t)LC) put(true);  // one pass: every entry is in it
            else put(false);
        }
        __syncthreads();
        for (uint32_t base = (uint32_t)wave * 64 * W; base < cnt; base += kThreads * W) {
            const uint32_t i0 = base + lane;
            if (ALL || i0 < cnt) {
                uint32_t e[W];
                bool valid[W];
#pragma unroll
                for (int w = 0; w < W; w++) {
                    const uint32_t iw = i0 + 64 * w;
                    valid[w] = iw < cnt;
                    e[w] = s_list[valid[w] ? iw : min(i0, cnt - 1)];  // an empty slot repeats a candidate
                }
                cands(t0, rd, e, valid);
            }
        }
        __syncthreads();
        flush_stage(st, t0, hkey, hmm, count, cap, t);
    }
}

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
                                                  stage_ref(s_stage), (strand << 15) | ls};
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

constexpr int kSeedListCap = 2048;  // smaller candidate list and hit stage: LDS room for the index
constexpr int kSeedStageCap = 128;

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
                                (strand << 15) | ls};
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

// N consecutive words at p (16-byte aligned for N = 4, 8-byte for N = 2).
template <int N>
__device__ __forceinline__ void load_words(const uint32_t *p, uint32_t (&m)[N]) {
    if constexpr (N == 4) {
        const uint4 v = *reinterpret_cast<const uint4 *>(p);
        m[0] = v.x;
        m[1] = v.y;
        m[2] = v.z;
        m[3] = v.w;
    } else if constexpr (N == 2) {
        const uint2 v = *reinterpret_cast<const uint2 *>(p);
        m[0] = v.x;
        m[1] = v.y;
    } else {
        m[0] = p[0];
    }
}

constexpr uint32_t kOr3 = 0xFE;  // a | b | c
constexpr uint32_t kMaj = 0xE8;  // at least two of a, b, c

// Bits set in at least one (any) and at least two (two) of the NP rows, in
// groups of three: a bit is in two rows when it is in two rows of one group
// or in two groups.
template <int NP, int NK, bool TWO>
__device__ __forceinline__ void combine_rows(const uint32_t (&m)[NP][NK], uint32_t (&any)[NK], uint32_t (&two)[NK]) {
    constexpr int G = (NP + 2) / 3;
    static_assert(G <= 3, "at most nine rows");
#pragma unroll
    for (int w = 0; w < NK; w++) {
        uint32_t g1[G], g2[G];
#pragma unroll
        for (int g = 0; g < G; g++) {
            const int n = NP - 3 * g < 3 ? NP - 3 * g : 3;
            if (n == 3) {
                g1[g] = __builtin_amdgcn_bitop3_b32(m[3 * g][w], m[3 * g + 1][w], m[3 * g + 2][w], kOr3);
                g2[g] = __builtin_amdgcn_bitop3_b32(m[3 * g][w], m[3 * g + 1][w], m[3 * g + 2][w], kMaj);
            } else if (n == 2) {
                g1[g] = m[3 * g][w] | m[3 * g + 1][w];
                g2[g] = m[3 * g][w] & m[3 * g + 1][w];
            } else {
                g1[g] = m[3 * g][w];
                g2[g] = 0;
            }
        }
        if constexpr (G == 1) {
            any[w] = g1[0];
            two[w] = g2[0];
        } else if constexpr (G == 2) {
            any[w] = g1[0] | g1[1];
            if (TWO) two[w] = __builtin_amdgcn_bitop3_b32(g1[0], g1[1], g2[0] | g2[1], kAndOr);
        } else {
            any[w] = __builtin_amdgcn_bitop3_b32(g1[0], g1[1], g1[2], kOr3);
            if (TWO)
                two[w] = __builtin_amdgcn_bitop3_b32(g2[0], g2[1], g2[2], kOr3) |
                         __builtin_amdgcn_bitop3_b32(g1[0], g1[1], g1[2], kMaj);
        }
    }
}

constexpr uint32_t kXor3 = 0x96;  // a ^ b ^ c

// Bits of word w set in at least `need` (uniform) of the NP rows: the
// bit-sliced count b3 b2 b1 b0 of the set rows from full adders (v_bitop3
// XOR3 and majority), then the comparison with need.
template <int NP, int CH>
__device__ __forceinline__ uint32_t at_least(const uint32_t (&m)[NP][CH], int w, uint32_t need) {
    static_assert(NP == 4 || NP == 6 || NP == 8, "row counts 4, 6, 8");
    auto fa = [](uint32_t x, uint32_t y, uint32_t z, uint32_t &sum, uint32_t &carry) {
        sum = __builtin_amdgcn_bitop3_b32(x, y, z, kXor3);
        carry = __builtin_amdgcn_bitop3_b32(x, y, z, kMaj);
    };
    uint32_t b0, b1, b2, b3 = 0, s1, c1;
    fa(m[0][w], m[1][w], m[2][w], s1, c1);
    if constexpr (NP == 4) {
        b0 = s1 ^ m[3][w];
        const uint32_t k = s1 & m[3][w];
        b1 = c1 ^ k;
        b2 = c1 & k;
    } else {
        uint32_t s2, c2;
        fa(m[3][w], m[4][w], m[5][w], s2, c2);
        if constexpr (NP == 6) {
            b0 = s1 ^ s2;
            fa(c1, c2, s1 & s2, b1, b2);
        } else {
            const uint32_t s3 = m[6][w] ^ m[7][w], c3 = m[6][w] & m[7][w];
            uint32_t k1, t0, t1;
            fa(s1, s2, s3, b0, k1);
            fa(c1, c2, c3, t0, t1);  // weights 2 and 4
            b1 = t0 ^ k1;
            const uint32_t k2 = t0 & k1;
            b2 = t1 ^ k2;
            b3 = t1 & k2;
        }
    }
    switch (need) {  // uniform: a scalar branch
        case 1: return b0 | __builtin_amdgcn_bitop3_b32(b1, b2, b3, kOr3);
        case 2: return __builtin_amdgcn_bitop3_b32(b1, b2, b3, kOr3);
        case 3: return __builtin_amdgcn_bitop3_b32(b1, b0, b2 | b3, kAndOr);
        case 4: return b2 | b3;
        case 5: return __builtin_amdgcn_bitop3_b32(b2, b1 | b0, b3, kAndOr);
        case 6: return __builtin_amdgcn_bitop3_b32(b2, b1, b3, kAndOr);
        case 7: return __builtin_amdgcn_bitop3_b32(b2, b1 & b0, b3, kAndOr);
        case 8: return b3;
        default: return 0u;
    }
}

// One-word window planes read from the staged tile (LDS word index and
// shift in 32-bit arithmetic); reverse complemented for strand 1.
__device__ __forceinline__ void window_lds(const uint32_t *s_lo, const uint32_t *s_hi, const uint32_t *s_ot,
                                           uint32_t bit, uint32_t strand, uint32_t &l, uint32_t &h, uint32_t &o) {
    const uint32_t wi = bit >> 5, sh = bit & 31;
    l = funnel(s_lo[wi + 1], s_lo[wi], sh);
    h = funnel(s_hi[wi + 1], s_hi[wi], sh);
    o = funnel(s_ot[wi + 1], s_ot[wi], sh);
    if (strand) {
        l = __builtin_bitreverse32(l);
        o = __builtin_bitreverse32(o);
        h = ~__builtin_bitreverse32(h) & ~o;  // complement: hi ^= 1 on bases
    }
}

// G tiles per workgroup (groups of kThreads threads) share one copy of the
// index in LDS; with G > 1 the candidate list holds kListCap entries.
// Register budget: six waves per SIMD, or three for wide rows (NK >=
// kMaskWideWords), whose LDS index allows no more waves anyway; with per-word
// counts their rows are read and counted for all words first, so that the
// reads of every chunk issue before the survivor loops.
constexpr int kMaskWideWords = 12;
template <int NK, int MINEX, int NP, int G>
__global__ __launch_bounds__(kThreads * G, NK >= kMaskWideWords ? 3 : 6) void fused_mask_kernel(MaskArgs a) {
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
            const uint32_t *row[NP];
#pragma unroll
            for (int p = 0; p < NP; p++) {
                const uint32_t f = md.first[p];  // bit | span << 8 | first row << 16
                const uint32_t a0 = f & 0xFF, span = (f >> 8) & 0xFF;
                const uint32_t code =
                    __builtin_amdgcn_ubfe(l, a0, span) | (__builtin_amdgcn_ubfe(h, a0, span) << span);
                row[p] = s_rows + (f >> 16) * NK + code * NK;  // the range's first row is uniform
            }
            auto count = [&](uint32_t r) {  // exact count of surviving record r
                const uint4 k = s_rec[r];
                const uint32_t y = __builtin_amdgcn_bitop3_b32(hc, k.y, oc, kXorOr);
                const uint32_t mm = (uint32_t)__popc(__builtin_amdgcn_bitop3_b32(lc, k.x, y, kXorOr));
                if ((int32_t)(k.z + mm) < 0) {  // the sink is built only on a hit
                    uint32_t ev = e[0];
                    asm volatile("" : "+v"(ev));  // keeps the sink's arithmetic on the (rare) hit path
                    emit_hit(TileSink{HitSink{a.hkey, a.hmm, a.count, a.cap, t0 + (ev >> 1), ev & 1, 0u},
                                      stage_ref(g_stage), ((ev & 1) << 15) | (ev >> 1)},
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
                    uint32_t m[NP][4];
#pragma unroll
                    for (int p = 0; p < NP; p++) load_words<4>(row[p] + 4 * c, m[p]);
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
                for (int p = 0; p < NP; p++) load_words<CH>(row[p] + CH * c, m[p]);
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

// Each wave queues the (candidate, guide) pairs whose family filter passed
// instead of branching into the exact count for one or two lanes at a time;
// a full queue (or the end of the candidate chunk) is drained one pair per
// lane.  An item is the guide record offset, the candidate entry | the mask
// of passing families << 16, and the pair's mismatch masks at the NS
// alignments (so that the drain needs neither the windows nor the guide's
// planes).
constexpr int kBulgeQueue = 64;

// QUEUE = false runs the exact counts in place, for the lanes that passed
// (fewer registers; cheaper when filters rarely pass).
template <int NW, bool GEN, int NS, bool QUEUE>
__global__ __launch_bounds__(kThreads, NW == 1 ? (QUEUE ? 5 : 6) : 1) void bulge_kernel(BulgeArgs a) {
    __shared__ uint32_t s_lo[kLdsWords], s_hi[kLdsWords], s_ot[kLdsWords];
    __shared__ uint16_t s_list[kListCap];
    __shared__ uint32_t s_wsum[4];
    __shared__ HitStageBuf<kStageCap> s_stage;
    __shared__ uint32_t s_qrec[4][QUEUE ? kBulgeQueue : 1], s_qinfo[4][QUEUE ? kBulgeQueue : 1];
    __shared__ uint32_t s_qm[4][QUEUE ? kBulgeQueue * NS * NW : 1];  // item i's masks at [i * NS * NW, ...)
    constexpr int KP = bulge_planes(NW, GEN), S = bulge_stride(NW, GEN, NS), FB = bulge_family_words(NW);
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t *qrec = s_qrec[wave], *qinfo = s_qinfo[wave], *qm = s_qm[wave];
    const kptr_t tab = (kptr_t)a.table;
    tile_candidates<1, true>(a.g, a.pd, a.limit, a.tile_stride, a.cand_count, s_lo, s_hi, s_ot, s_list, s_wsum,
                             stage_ref(s_stage), a.hkey, a.hmm, a.count, a.cap,
                             [&](uint64_t t0, auto rd, const uint32_t (&e)[1], const bool (&valid)[1]) {
        // exact counts of the queued pairs, one per lane (records read per lane)
        auto drain = [&](uint32_t qn) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (lane < qn) {
                const uint32_t info = qinfo[lane];
                const kptr_t kp = tab + qrec[lane];
                const uint32_t ce = info & 0xFFFFu, fam = info >> 16;
                uint32_t M[NS][NW];
#pragma unroll
                for (int u = 0; u < NS; u++)
#pragma unroll
                    for (int w = 0; w < NW; w++) M[u][w] = qm[lane * NS * NW + u * NW + w];
                const kptr_t vb = tab + kp[KP];
                const TileSink sink{HitSink{a.hkey, a.hmm, a.count, a.cap, t0 + (ce >> 1), ce & 1, vb[0]},
                                    stage_ref(s_stage), ((ce & 1) << 15) | (ce >> 1)};
#pragma unroll
                for (int u = 0; u < NS; u++) {
                    if ((fam >> u) & 1) {
                        const bool lft0 = u == 0 || a.rev;
                        uint32_t left[NW], right[NW];
#pragma unroll
                        for (int w = 0; w < NW; w++) {
                            left[w] = lft0 ? M[0][w] : M[u][w];
                            right[w] = lft0 ? M[u][w] : M[0][w];
                        }
                        bulge_family<NW>(tab, vb + 4 + FB * u, left, right, a.gap[u], kp[KP + 1 + u], sink);
                    }
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
        };

        uint32_t l[NS][NW], h[NS][NW], o[NS][NW];
        bulge_windows<NW, NS>(a, rd, e[0], l, h, o);
        uint32_t cls[GEN ? NS : 1][5][NW];
        if (GEN) {
#pragma unroll
            for (int u = 0; u < NS; u++) class_planes<NW>(l[u], h[u], o[u], cls[u]);
        }
        uint32_t qn = 0;  // wave-uniform queue length
        kptr_t kp = tab;
        for (uint32_t c = 0; c < a.nunits; c++) {
            uint32_t lc[NS][NW], hc[NS][NW], oc[NS][NW];
            if (!GEN) {
#pragma unroll
                for (int w = 0; w < NW; w++) {
                    const uint32_t care = kp[w];
#pragma unroll
                    for (int u = 0; u < NS; u++) {
                        lc[u][w] = l[u][w] & care;
                        hc[u][w] = h[u][w] & care;
                        oc[u][w] = o[u][w] & care;
                    }
                }
            }
            const uint32_t n = kp[NW];
            kp += kClassHdr;
            for (uint32_t r = 0; r < n; r++, kp += S) {
                uint32_t M[NS][NW];
#pragma unroll
                for (int u = 0; u < NS; u++)
#pragma unroll
                    for (int w = 0; w < NW; w++) {
                        if (GEN) {
                            M[u][w] = general_mask<NW>(kp, cls[u], w);
                        } else {
                            const uint32_t x = __builtin_amdgcn_bitop3_b32(hc[u][w], kp[NW + w], oc[u][w], kXorOr);
                            M[u][w] = __builtin_amdgcn_bitop3_b32(lc[u][w], kp[w], x, kXorOr);
                        }
                    }
                int32_t f[NS];
                {
                    uint32_t s = kp[KP + 1];
#pragma unroll
                    for (int w = 0; w < NW; w++) s += (uint32_t)__popc(M[0][w]);
                    f[0] = (int32_t)s;
                }
                int32_t acc = f[0];
#pragma unroll
                for (int u = 1; u < NS; u++) {
                    uint32_t s = kp[KP + 1 + u];
#pragma unroll
                    for (int w = 0; w < NW; w++) s += (uint32_t)__popc(M[u][w] & M[0][w]);
                    f[u] = (int32_t)s;
                    acc |= f[u];
                }
                const bool pass = valid[0] && acc < 0;
                if (!QUEUE) {
                    if (pass) {
                        const kptr_t vb = tab + kp[KP];
                        const uint32_t ce = e[0];
                        const TileSink sink{HitSink{a.hkey, a.hmm, a.count, a.cap, t0 + (ce >> 1), ce & 1, vb[0]},
                                            stage_ref(s_stage), ((ce & 1) << 15) | (ce >> 1)};
#pragma unroll
                        for (int u = 0; u < NS; u++) {
                            if (f[u] < 0) {
                                const bool lft0 = u == 0 || a.rev;
                                uint32_t left[NW], right[NW];
#pragma unroll
                                for (int w = 0; w < NW; w++) {
                                    left[w] = lft0 ? M[0][w] : M[u][w];
                                    right[w] = lft0 ? M[u][w] : M[0][w];
                                }
                                bulge_family<NW>(tab, vb + 4 + FB * u, left, right, a.gap[u], kp[KP + 1 + u], sink);
                            }
                        }
                    }
                    continue;
                }
                const uint64_t m = __ballot(pass);
                if (m) {
                    const uint32_t np = (uint32_t)__popcll(m);
                    if (qn + np > (uint32_t)kBulgeQueue) {
                        drain(qn);
                        qn = 0;
                    }
                    if (pass) {
                        uint32_t fam = 0;
#pragma unroll
                        for (int u = 0; u < NS; u++) fam |= (f[u] < 0 ? 1u : 0u) << u;
                        const uint32_t slot =
                            qn + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                           __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
                        qrec[slot] = (uint32_t)(kp - tab);
                        qinfo[slot] = e[0] | (fam << 16);
#pragma unroll
                        for (int u = 0; u < NS; u++)
#pragma unroll
                            for (int w = 0; w < NW; w++) qm[slot * NS * NW + u * NW + w] = M[u][w];
                    }
                    qn += np;
                }
            }
        }
        if (QUEUE && qn) drain(qn);
    });
}

// bulge_index_kernel: bulge_kernel's families behind a bit-set seed index of
// the guides (one care class, A/C/G/T guides, one-word windows).  A variant
// of split j and RNA gap g compares the guide positions below j at one window
// alignment and those from j + g at another, so it breaks at most max(1, R)
// of the index's ranges and, within threshold t, matches at least
// parts - max(1, R) - t of them exactly at alignment 0 or at the family's
// alignment u.  Per candidate each range's row is read at every alignment;
// per family u the rows at 0 and u are OR-ed and a guide survives when, for
// some family, at least its word's need of them hold it.  Only survivors get
// the family filters and exact counts of bulge_kernel (in place).  Index (LDS): rows of nk words per (range, code),
// then per slot [glo, ghi, variant block offset, family filter bias per
// alignment] (kBulgeSlotWords words).
struct BulgeIndexArgs {
    BulgeArgs b;
    MaskDesc md;
    const uint32_t *idx;
};

template <int NK, int NS, int NP>
__global__ __launch_bounds__(kThreads, NK == 4 ? 3 : 5) void bulge_index_kernel(BulgeIndexArgs ia) {
    __shared__ uint32_t s_lo[kLdsWords], s_hi[kLdsWords], s_ot[kLdsWords];
    __shared__ uint16_t s_list[kListCap];
    __shared__ uint32_t s_wsum[4];
    __shared__ HitStageBuf<kStageCap> s_stage;
    extern __shared__ __align__(16) uint32_t s_idx[];
    const BulgeArgs &a = ia.b;
    const MaskDesc &md = ia.md;
    for (uint32_t i = threadIdx.x; i < md.nwords / 4; i += kThreads)
        reinterpret_cast<uint4 *>(s_idx)[i] = reinterpret_cast<const uint4 *>(ia.idx)[i];
    // visible after tile_candidates' first barrier
    constexpr int FB = bulge_family_words(1);
    const kptr_t tab = (kptr_t)a.table;
    const uint32_t care = md.care;
    const uint4 *s_slot = reinterpret_cast<const uint4 *>(s_idx + md.recw);
    tile_candidates<1, false>(a.g, a.pd, a.limit, a.tile_stride, a.cand_count, s_lo, s_hi, s_ot, s_list, s_wsum,
                              stage_ref(s_stage), a.hkey, a.hmm, a.count, a.cap,
                              [&](uint64_t t0, auto rd, const uint32_t (&e)[1], const bool (&)[1]) {
        uint32_t l[NS][1], h[NS][1], o[NS][1];
        bulge_windows<1, NS>(a, rd, e[0], l, h, o);
        // row of range p at alignment u
        auto row = [&](int p, int u, uint32_t (&r)[NK]) {
            const uint32_t f = md.first[p];  // bit | span << 8 | first row << 16
            const uint32_t a0 = f & 0xFF, span = (f >> 8) & 0xFF;
            const uint32_t code =
                __builtin_amdgcn_ubfe(l[u][0], a0, span) | (__builtin_amdgcn_ubfe(h[u][0], a0, span) << span);
            load_words<NK>(s_idx + (f >> 16) * NK + code * NK, r);  // padding parts: span 0, the zero row
        };
        // a variant of family u compares every unbroken range at alignment 0
        // or u: per family, the ranges exact at either are counted
        uint32_t m0[NP][NK], sel[NK];
#pragma unroll
        for (int p = 0; p < NP; p++) row(p, 0, m0[p]);
#pragma unroll
        for (int w = 0; w < NK; w++) sel[w] = 0;
#pragma unroll
        for (int u = 1; u < NS; u++) {
            uint32_t m[NP][NK];
#pragma unroll
            for (int p = 0; p < NP; p++) {
                row(p, u, m[p]);
#pragma unroll
                for (int k = 0; k < NK; k++) m[p][k] |= m0[p][k];
            }
#pragma unroll
            for (int w = 0; w < NK; w++) sel[w] |= at_least<NP, NK>(m, w, (md.need[w / 8] >> (4 * (w % 8))) & 15u);
        }
        uint32_t lc[NS], hc[NS], oc[NS];
#pragma unroll
        for (int u = 0; u < NS; u++) {
            lc[u] = l[u][0] & care;
            hc[u] = h[u][0] & care;
            oc[u] = o[u][0] & care;
        }
        const uint32_t ce = e[0];
#pragma unroll
        for (int w = 0; w < NK; w++) {
            for (uint32_t x = sel[w]; x; x &= x - 1) {
                const uint32_t slot = 32 * w + (uint32_t)__builtin_ctz(x);
                const uint4 s0 = s_slot[2 * slot], s1 = s_slot[2 * slot + 1];
                const uint32_t bias[5] = {s0.w, s1.x, s1.y, s1.z, s1.w};
                uint32_t M[NS][1];
                int32_t fl[NS];
                int32_t acc = 0;
#pragma unroll
                for (int u = 0; u < NS; u++) {
                    const uint32_t x = __builtin_amdgcn_bitop3_b32(hc[u], s0.y, oc[u], kXorOr);
                    M[u][0] = __builtin_amdgcn_bitop3_b32(lc[u], s0.x, x, kXorOr);
                }
#pragma unroll
                for (int u = 0; u < NS; u++) {
                    fl[u] = (int32_t)(bias[u] + (uint32_t)__popc(u == 0 ? M[0][0] : M[u][0] & M[0][0]));
                    acc |= fl[u];
                }
                if (acc >= 0) continue;
                const kptr_t vb = tab + s0.z;
                const TileSink sink{HitSink{a.hkey, a.hmm, a.count, a.cap, t0 + (ce >> 1), ce & 1, vb[0]},
                                    stage_ref(s_stage), ((ce & 1) << 15) | (ce >> 1)};
#pragma unroll
                for (int u = 0; u < NS; u++) {
                    if (fl[u] < 0) {
                        const bool lft0 = u == 0 || a.rev;
                        const uint32_t left[1] = {lft0 ? M[0][0] : M[u][0]}, right[1] = {lft0 ? M[u][0] : M[0][0]};
                        bulge_family<1>(tab, vb + 4 + FB * u, left, right, a.gap[u], bias[u], sink);
                    }
                }
            }
        }
    });
}

// ------------------------------------------------------------------ launchers

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

// Trial sampling: tile_stride > 1 launches one workgroup per
// tile_stride-th tile, at most max_tiles of them.
static dim3 tile_grid(uint64_t limit, uint32_t tile_stride, uint32_t max_tiles) {
    const uint64_t ntiles = (limit + kTilePos - 1) / kTilePos;
    uint64_t n = (ntiles + tile_stride - 1) / tile_stride;
    if (max_tiles) n = std::min<uint64_t>(n, max_tiles);
    return dim3((uint32_t)n);
}

template <int NW, bool GEN>
static void fused_variant(int variant, dim3 grid, const FusedArgs &a, hipStream_t s) {
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
    hipLaunchKernelGGL(fused_seed_kernel, grid, dim3(kThreads), (lds_bytes + 15) / 16 * 16, s, a);
    return hipGetLastError();
}

template <int NK, int MINEX, int NP, int G>
static hipError_t mask_launch(const MaskArgs &a, uint32_t lds, hipStream_t s) {
    const dim3 grid((a.ntile + G - 1) / G);
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
    if (a.md.minex == 2) return mask_groups_launch<NK, 2, NP>(a, lds, s);
    if (a.md.minex == 3) return mask_groups_launch<NK, 3, NP>(a, lds, s);
    if constexpr (NK <= 8) return mask_groups_launch<NK, 1, NP>(a, lds, s);  // wider tables: minex 2, 3 only
    return hipErrorInvalidValue;
}

template <int NK>
static hipError_t mask_parts(const MaskArgs &a, uint32_t lds, hipStream_t s) {
    if (a.md.parts <= 4) return mask_minex<NK, 4>(a, lds, s);
    if (a.md.parts <= 6) return mask_minex<NK, 6>(a, lds, s);
    return mask_minex<NK, 8>(a, lds, s);
}

hipError_t launch_mask(const DevGenome &g, const PatternDesc &pd, const MaskDesc &md, uint32_t cbase,
                       const uint32_t *idx, uint64_t *hkey, uint16_t *hmm, unsigned long long *count,
                       unsigned long long *cand_count, uint64_t cap, hipStream_t s, uint32_t tile_stride,
                       uint32_t max_tiles) {
    const uint64_t limit = g.total >= pd.P ? g.total - pd.P + 1 : 0;
    if (limit == 0 || md.nrec == 0) return hipSuccess;
    if (tile_stride == 0 || md.parts == 0 || md.parts > (uint32_t)kSeedMaxParts || md.minex < 1 || md.minex > 3 ||
        md.nrec > 32 * md.nk || md.nwords % 4 || md.recw % 4 || md.recw + 4 * md.nrec != md.nwords ||
        (uintptr_t)idx % 16)
        return hipErrorInvalidValue;
    for (uint32_t p = 0; p < (uint32_t)kSeedMaxParts; p++) {
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

template <int NW, bool GEN, int NS>
static void bulge_variant(int variant, dim3 grid, const BulgeArgs &a, hipStream_t s) {
    if (variant == 1) hipLaunchKernelGGL((bulge_kernel<NW, GEN, NS, false>), grid, dim3(kThreads), 0, s, a);
    else hipLaunchKernelGGL((bulge_kernel<NW, GEN, NS, true>), grid, dim3(kThreads), 0, s, a);
}

template <int NW, bool GEN>
static hipError_t launch_bulge_ns(int ns, int variant, dim3 grid, const BulgeArgs &a, hipStream_t s) {
    switch (ns) {
        case 2: bulge_variant<NW, GEN, 2>(variant, grid, a, s); break;
        case 3: bulge_variant<NW, GEN, 3>(variant, grid, a, s); break;
        case 4: bulge_variant<NW, GEN, 4>(variant, grid, a, s); break;
        case 5: bulge_variant<NW, GEN, 5>(variant, grid, a, s); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_bulge(const DevGenome &g, const PatternDesc &pd, int nw, bool gen, int ns, const int32_t *cb,
                        const uint32_t *gap, bool rev, const uint32_t *table, uint32_t nunits, uint64_t *hkey,
                        uint16_t *hmm, unsigned long long *count, unsigned long long *cand_count, uint64_t cap,
                        hipStream_t s, int variant, uint32_t tile_stride, uint32_t max_tiles) {
    const uint64_t limit = g.total >= pd.P ? g.total - pd.P + 1 : 0;
    if (limit == 0 || nunits == 0) return hipSuccess;
    if (ns < 2 || ns > kMaxShifts || (nw != 1 && nw != 2) || tile_stride == 0) return hipErrorInvalidValue;
    BulgeArgs a{};
    a.g = g;
    a.pd = pd;
    for (int u = 0; u < ns; u++) {
        a.cb[u] = cb[u];
        a.gap[u] = gap[u];
    }
    a.rev = rev ? 1u : 0u;
    a.nunits = nunits;
    a.tile_stride = tile_stride;
    a.table = table;
    a.limit = limit;
    a.hkey = hkey;
    a.hmm = hmm;
    a.count = count;
    a.cand_count = cand_count;
    a.cap = cap;
    const dim3 grid = tile_grid(limit, tile_stride, max_tiles);
    if (nw == 1)
        return gen ? launch_bulge_ns<1, true>(ns, variant, grid, a, s)
                   : launch_bulge_ns<1, false>(ns, variant, grid, a, s);
    return gen ? launch_bulge_ns<2, true>(ns, variant, grid, a, s) : launch_bulge_ns<2, false>(ns, variant, grid, a, s);
}

template <int NK, int NS>
static hipError_t bulge_index_parts(uint32_t parts, dim3 grid, const BulgeIndexArgs &a, uint32_t lds, hipStream_t s) {
    if (parts <= 4) hipLaunchKernelGGL((bulge_index_kernel<NK, NS, 4>), grid, dim3(kThreads), lds, s, a);
    else if (parts <= 6) hipLaunchKernelGGL((bulge_index_kernel<NK, NS, 6>), grid, dim3(kThreads), lds, s, a);
    else hipLaunchKernelGGL((bulge_index_kernel<NK, NS, 8>), grid, dim3(kThreads), lds, s, a);
    return hipGetLastError();
}

template <int NK>
static hipError_t bulge_index_ns(int ns, uint32_t parts, dim3 grid, const BulgeIndexArgs &a, uint32_t lds,
                                 hipStream_t s) {
    switch (ns) {
        case 2: return bulge_index_parts<NK, 2>(parts, grid, a, lds, s);
        case 3: return bulge_index_parts<NK, 3>(parts, grid, a, lds, s);
        case 4: return bulge_index_parts<NK, 4>(parts, grid, a, lds, s);
        case 5: return bulge_index_parts<NK, 5>(parts, grid, a, lds, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_bulge_index(const DevGenome &g, const PatternDesc &pd, int ns, const int32_t *cb,
                              const uint32_t *gap, bool rev, const uint32_t *table, const MaskDesc &md,
                              const uint32_t *idx, uint64_t *hkey, uint16_t *hmm, unsigned long long *count,
                              unsigned long long *cand_count, uint64_t cap, hipStream_t s, uint32_t tile_stride,
                              uint32_t max_tiles) {
    const uint64_t limit = g.total >= pd.P ? g.total - pd.P + 1 : 0;
    if (limit == 0 || md.nrec == 0) return hipSuccess;
    // the kernel reads md.first[0..7], the rows they address, nrec slots
    // after recw, and need nibbles of nk words
    if (ns < 2 || ns > kMaxShifts || tile_stride == 0 || md.parts == 0 || md.parts > (uint32_t)kSeedMaxParts ||
        md.minex != 3 || (md.nk != 1 && md.nk != 2 && md.nk != 4) || md.nrec > 32 * md.nk || md.recw % 4 ||
        md.nwords % 4 || md.recw + kBulgeSlotWords * md.nrec != md.nwords || md.nwords * 4 > kBulgeIndexMaxBytes ||
        (uintptr_t)idx % 16)
        return hipErrorInvalidValue;
    for (uint32_t p = 0; p < (uint32_t)kSeedMaxParts; p++) {
        const uint32_t span = (md.first[p] >> 8) & 0xFF, rbase = md.first[p] >> 16;
        if (p >= md.parts ? md.first[p] != md.zrow << 16
                          : span == 0 || span > (uint32_t)kMaskMaxSpan || (md.first[p] & 0xFF) + span > 32 ||
                                rbase + (1u << (2 * span)) > md.zrow)
            return hipErrorInvalidValue;
    }
    if ((uint64_t)(md.zrow + 1) * md.nk > md.recw) return hipErrorInvalidValue;
    BulgeIndexArgs a{};
    BulgeArgs &b = a.b;
    b.g = g;
    b.pd = pd;
    for (int u = 0; u < ns; u++) {
        b.cb[u] = cb[u];
        b.gap[u] = gap[u];
    }
    b.rev = rev ? 1u : 0u;
    b.nunits = 1;
    b.tile_stride = tile_stride;
    b.table = table;
    b.limit = limit;
    b.hkey = hkey;
    b.hmm = hmm;
    b.count = count;
    b.cand_count = cand_count;
    b.cap = cap;
    a.md = md;
    a.idx = idx;
    const dim3 grid = tile_grid(limit, tile_stride, max_tiles);
    const uint32_t lds = md.nwords * 4;
    switch (md.nk) {
        case 1: return bulge_index_ns<1>(ns, md.parts, grid, a, lds, s);
        case 2: return bulge_index_ns<2>(ns, md.parts, grid, a, lds, s);
        default: return bulge_index_ns<4>(ns, md.parts, grid, a, lds, s);
    }
}

// Radix sort of the hit records by their 64-bit key (bits [0, end_bit)).
// With tmp == nullptr only *tmp_bytes is set.
hipError_t sort_hits(void *tmp, size_t *tmp_bytes, const uint64_t *kin, uint64_t *kout, const uint16_t *vin,
                     uint16_t *vout, uint64_t n, unsigned end_bit, hipStream_t s) {
    return rocprim::radix_sort_pairs(tmp, *tmp_bytes, kin, kout, vin, vout, (size_t)n, 0u, end_bit, s);
}

}  // namespace cof
