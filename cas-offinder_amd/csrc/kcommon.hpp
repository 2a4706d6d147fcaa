// Device helpers shared by the gfx950 kernels of the search path.
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
// Mismatch counting works on NW 32-bit words of the window planes taken at
// key position cbase.  A key whose comparable bases are all A/C/G/T is
// (klo, khi, care): mismatches = popcount(((lo^klo) | (hi^khi) | ot) & care);
// a key with IUPAC codes is five class masks (the classes that mismatch at
// each position): popcount(OR_c class_c & mask_c).  The threshold is folded in
// as a bias -(threshold+1) added by the popcount, so a score < 0 is a hit.
// The '-' strand reverse complements the window (bit reverse, hi ^= 1) and
// compares the forward key, which counts the same mismatches as the reverse
// complemented key against the forward text (cas-offinder.cl:123-154).
//
// Kernels, one translation unit per family:
//   k_plain.hip        finder_count_kernel, finder_write_kernel, comparer_kernel
//                      (the reference's launch model)
//   k_fused.hip        fused_kernel, fused_seed_kernel
//   k_mask.hip         fused_mask_kernel (bit-set seed index)
//   k_bulge.hip        bulge_kernel (device-side bulge expansion)
//   k_bulge_index.hip  bulge_index_kernel
//   k_sort.hip         device radix sort of the hits
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "devtypes.hpp"

namespace cof {

// The kernel instance that the calling thread launched last, named as
// profilers print it ("cof::bulge_kernel<1, false, 2, true>"); the engine
// reads it after each batched launch (last_launched_kernel).
inline thread_local std::string g_kernel_name;
inline std::string kname_arg(int v) { return std::to_string(v); }
inline std::string kname_arg(bool v) { return v ? "true" : "false"; }
template <typename... A>
inline void note_kernel(const char *name, A... a) {
    std::string s = std::string("cof::") + name;
    if constexpr (sizeof...(A) > 0) {
        const std::string args[] = {kname_arg(a)...};
        s += "<";
        for (size_t i = 0; i < sizeof...(A); i++) s += (i ? ", " : "") + args[i];
        s += ">";
    }
    g_kernel_name = std::move(s);
}

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
// channel).  An entry packs global key index (bits 24..47), mismatches
// (17..23), strand (16) and the window start's offset from the tile start
// plus kStageBias (0..15): a bulge guide whose padding precedes the window
// reports a start up to kMaxShifts - 1 positions before the tile.
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
constexpr int32_t kStageBias = 16;
static_assert(kTilePos + kStageBias <= (1 << 16) && kMaxShifts <= kStageBias && kMaxPattern < (1 << 7),
              "stage entry fields");

// Low bits of a stage entry: strand and the window start's offset from the
// tile start (at least -kStageBias).
__device__ __forceinline__ uint32_t stage_low(uint32_t strand, int32_t off) {
    return strand << 16 | (uint32_t)(off + kStageBias);
}

struct TileSink {
    HitSink g;      // direct path when the stage is full
    StageRef st;
    uint32_t low;   // stage_low(strand, window start - tile start)
};

__device__ __forceinline__ void emit_hit(const TileSink &o, uint32_t key, uint32_t mm) {
    const uint32_t i = atomicAdd(&o.st.h->n, 1u);
    if (i < o.st.cap)
        o.st.ent[i] = ((uint64_t)(o.g.gbase + key) << 24) | (mm << 17) | o.low;
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
            hkey[o] = ((e >> 24) << kHitKeyShift) | ((t0 + (e & 0xFFFF) - kStageBias) << 1) | ((e >> 16) & 1);
            hmm[o] = (uint16_t)((e >> 17) & 0x7F);
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

constexpr int kListCap = 4096;  // LDS candidate entries per pass

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
// fin(t0, rd) runs on every thread after the candidates of each pass and
// before the stage is flushed (per-wave queues drain there).  pre() runs on
// every thread while the tile's plane loads are in flight (before the
// first barrier: LDS it writes is visible to the candidates).
struct NoFin {
    template <typename Rd>
    __device__ void operator()(uint64_t, Rd) const {}
};
struct NoPre {
    __device__ void operator()() const {}
};
template <int W, bool ALL = false, int LC = kListCap, int G = 1, typename Cands, typename Fin = NoFin,
          typename Pre = NoPre>
__device__ __forceinline__ void tile_candidates(const DevGenome &g, const PatternDesc &pd, uint64_t limit,
                                                uint32_t tile_stride, unsigned long long *cand_count,
                                                uint32_t *s_lo, uint32_t *s_hi, uint32_t *s_ot, uint16_t *s_list,
                                                uint32_t *s_wsum, StageRef st, uint64_t *hkey, uint16_t *hmm,
                                                unsigned long long *count, uint64_t cap, Cands &&cands,
                                                uint32_t ntile = 0, uint32_t *s_nmax = nullptr, Fin &&fin = Fin{},
                                                Pre &&pre = Pre{}) {
    const int t = G == 1 ? (int)threadIdx.x : (int)(threadIdx.x & (kThreads - 1));
    const uint32_t grp = G == 1 ? 0u : threadIdx.x / kThreads;
    const uint32_t gi = G == 1 ? blockIdx.x : blockIdx.x * G + grp;
    const bool live = G == 1 || gi < ntile;
    const uint32_t tile = live ? gi * tile_stride : 0u;
    const int64_t w0 = (int64_t)tile * kTileWords;
    const uint64_t ppos = g.tile_ppos[tile];
    {
        TileRegs r;
        load_tile(g, w0, r, t);
        pre();
        store_tile(r, s_lo, s_hi, s_ot, t);
    }
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
        fin(t0, rd);
        __syncthreads();
        flush_stage(st, t0, hkey, hmm, count, cap, t);
    }
}

constexpr int kSeedListCap = 2048;  // smaller candidate list and hit stage: LDS room for the index
constexpr int kSeedStageCap = 128;

namespace {

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
    static_assert(NP == 4 || NP == 6 || NP == 8 || NP == 10 || NP == 12, "row counts 4, 6, 8, 10, 12");
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
        } else if constexpr (NP == 8) {
            const uint32_t s3 = m[6][w] ^ m[7][w], c3 = m[6][w] & m[7][w];
            uint32_t k1, t0, t1;
            fa(s1, s2, s3, b0, k1);
            fa(c1, c2, c3, t0, t1);  // weights 2 and 4
            b1 = t0 ^ k1;
            const uint32_t k2 = t0 & k1;
            b2 = t1 ^ k2;
            b3 = t1 & k2;
        } else if constexpr (NP == 10) {
            uint32_t s3, c3, x, k1, t0, t1, t2;
            fa(m[6][w], m[7][w], m[8][w], s3, c3);
            fa(s1, s2, s3, x, k1);  // weight 1: x, m9
            b0 = x ^ m[9][w];
            const uint32_t k2 = x & m[9][w];
            fa(c1, c2, c3, t0, t1);  // weight 2: t0, k1, k2
            fa(t0, k1, k2, b1, t2);
            b2 = t1 ^ t2;  // weight 4: t1, t2 (at most 10 rows: no carry past weight 8)
            b3 = t1 & t2;
        } else {
            uint32_t s3, c3, s4, c4, x, k1, y1, d1, y2, d2;
            fa(m[6][w], m[7][w], m[8][w], s3, c3);
            fa(m[9][w], m[10][w], m[11][w], s4, c4);
            fa(s1, s2, s3, x, k1);  // weight 1: x, s4
            b0 = x ^ s4;
            const uint32_t k2 = x & s4;
            fa(c1, c2, c3, y1, d1);  // weight 2: y1, c4, k1, k2
            fa(c4, k1, k2, y2, d2);
            b1 = y1 ^ y2;
            fa(d1, d2, y1 & y2, b2, b3);  // weight 4 (at most 12 rows: b3 is weight 8)
        }
    }
    // count >= need, bit-sliced from the low bit up, without branches: with
    // n_i = bit i of the uniform need as an all-ones / all-zero scalar,
    // ge_0 = b0 | ~n0 and ge_i = (b_i & ~n_i) | (~(b_i ^ n_i) & ge_{i-1}),
    // one v_bitop3 each (truth tables over b_i * 4 + n_i * 2 + ge_{i-1}).
    // A need above the largest count (15: a word without records) keeps nothing.
    // The empty asm keeps the masks from being hoisted out of the candidate
    // loop (four per word of a wide table held live would spill): they cost
    // one scalar bit-field extract each where they are used.
    asm volatile("" : "+s"(need));
    auto nbit = [&](int i) { return (uint32_t)((int32_t)(need << (31 - i)) >> 31); };
    constexpr uint32_t kGe0 = 0xF3, kGe = 0xB2;
    uint32_t ge = __builtin_amdgcn_bitop3_b32(b0, nbit(0), 0u, kGe0);
    ge = __builtin_amdgcn_bitop3_b32(b1, nbit(1), ge, kGe);
    ge = __builtin_amdgcn_bitop3_b32(b2, nbit(2), ge, kGe);
    return __builtin_amdgcn_bitop3_b32(b3, nbit(3), ge, kGe);
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

}  // namespace

// Trial sampling: tile_stride > 1 launches one workgroup per
// tile_stride-th tile, at most max_tiles of them.
inline dim3 tile_grid(uint64_t limit, uint32_t tile_stride, uint32_t max_tiles) {
    const uint64_t ntiles = (limit + kTilePos - 1) / kTilePos;
    uint64_t n = (ntiles + tile_stride - 1) / tile_stride;
    if (max_tiles) n = std::min<uint64_t>(n, max_tiles);
    return dim3((uint32_t)n);
}

}  // namespace cof
