// gfx950 kernels of the search path.
//
// Genome planes (lo, hi, ot) are 32-position words; a 256-thread workgroup
// owns one tile of 256 words (8192 window starts) and stages the tile plus
// kHaloL / kHaloR words of halo in LDS.  Each thread tests the PAM of the 32
// starts of its word on both strands with bitwise class masks
// (finder_masks), so one lane does 32 positions of cas-offinder.cl:26-63 at
// once.  Windows holding a ';' (poison list) or running past the genome end
// are cleared (cas-offinder.cl:64-67).
//
//   finder_kernel    plain path: compacts (locus, flag) like cas-offinder.cl:68-70
//   comparer_kernel  plain path: one compare key per launch over the
//                    candidate list (cas-offinder.cl:73-155)
//   fused_kernel     batched path: finder + compaction into LDS + every
//                    compare key of the query per candidate, no candidate
//                    list in HBM
//
// Mismatch counting: a window of NW words per plane is turned into five class
// planes (T, C, A, G, other); a key is five class masks (the classes that
// mismatch at each key position), so the count is
// popcount(OR_c class_c & mask_c).  The '-' strand is done by reverse
// complementing the window (bit reverse, hi ^= 1) and comparing the forward
// key, which counts the same mismatches as the reverse-complemented key
// against the forward text (cas-offinder.cl:123-154).
#include <hip/hip_runtime.h>

#include "devtypes.hpp"

namespace cof {

namespace {

__device__ __forceinline__ uint32_t funnel(uint32_t hw, uint32_t lw, uint32_t sh) {
    return __builtin_amdgcn_alignbit(hw, lw, sh);  // bits [sh, sh+32) of hw:lw
}

__device__ __forceinline__ uint32_t sel_mask(uint32_t m, uint32_t c) { return ((m >> c) & 1u) ? ~0u : 0u; }

// planes -> class-match bits for a 5-bit class set `m` (o positions have lo = hi = 0)
__device__ __forceinline__ uint32_t class_match(uint32_t m, uint32_t l, uint32_t h, uint32_t o) {
    const uint32_t T = ~(l | h | o), C = l & ~h, A = h & ~l, G = l & h;
    return (T & sel_mask(m, kClsT)) | (C & sel_mask(m, kClsC)) | (A & sel_mask(m, kClsA)) |
           (G & sel_mask(m, kClsG)) | (o & sel_mask(m, kClsO));
}

// PAM test of the 32 starts of LDS word i0 (and words i0+1, i0+2 for the
// pattern reach): bit b of *fwd / *rev = start b passes the forward /
// reverse-complement pattern.
__device__ __forceinline__ void finder_masks(const PatternDesc &pd, const uint32_t *s_lo, const uint32_t *s_hi,
                                             const uint32_t *s_ot, int i0, uint32_t *fwd, uint32_t *rev) {
    uint32_t L0 = s_lo[i0], L1 = s_lo[i0 + 1], L2 = s_lo[i0 + 2];
    uint32_t H0 = s_hi[i0], H1 = s_hi[i0 + 1], H2 = s_hi[i0 + 2];
    uint32_t O0 = s_ot[i0], O1 = s_ot[i0 + 1], O2 = s_ot[i0 + 2];
    uint32_t f = ~0u, r = ~0u;
    for (uint32_t i = 0; i < pd.nf; i++) {
        const uint32_t k = pd.fk[i], sh = k & 31;
        const bool up = k >= 32;
        const uint32_t l = funnel(up ? L2 : L1, up ? L1 : L0, sh);
        const uint32_t h = funnel(up ? H2 : H1, up ? H1 : H0, sh);
        const uint32_t o = funnel(up ? O2 : O1, up ? O1 : O0, sh);
        f &= class_match(pd.fm[i], l, h, o);
    }
    for (uint32_t i = 0; i < pd.nr; i++) {
        const uint32_t k = pd.rk[i], sh = k & 31;
        const bool up = k >= 32;
        const uint32_t l = funnel(up ? L2 : L1, up ? L1 : L0, sh);
        const uint32_t h = funnel(up ? H2 : H1, up ? H1 : H0, sh);
        const uint32_t o = funnel(up ? O2 : O1, up ? O1 : O0, sh);
        r &= class_match(pd.rm[i], l, h, o);
    }
    *fwd = f;
    *rev = r;
}

// Bits of the starts [s0, s0+32) that are inside [lo, hi): window starts
// allowed by the genome end and the slab.
__device__ __forceinline__ uint32_t range_mask(uint64_t s0, uint64_t lo, uint64_t hi) {
    uint32_t m = ~0u;
    if (s0 + 32 > hi) m = s0 >= hi ? 0u : (uint32_t)((1ull << (hi - s0)) - 1);
    if (s0 < lo) m &= lo >= s0 + 32 ? 0u : ~(uint32_t)((1ull << (lo - s0)) - 1);
    return m;
}

// Clears the starts of word [s0, s0+32) whose window [s, s+P) holds a poison
// position.  t0 / t1: the start range the workgroup covers.
__device__ __forceinline__ uint32_t poison_mask(const DevGenome &g, uint32_t P, uint64_t t0, uint64_t t1,
                                                uint64_t s0) {
    // first poison >= t0 (uniform binary search)
    uint64_t lo = 0, n = g.npoison;
    while (n > 0) {
        const uint64_t half = n >> 1;
        if (g.poison[lo + half] < t0) { lo += half + 1; n -= half + 1; }
        else n = half;
    }
    uint32_t keep = ~0u;
    for (uint64_t i = lo;; i++) {
        const uint64_t x = g.poison[i];  // poison[npoison] is the sentinel
        if (x >= t1 + P - 1) break;
        const uint64_t a = x + 1 >= P ? x + 1 - P : 0;  // starts [a, x] see x
        if (x >= s0 && a < s0 + 32) {
            const uint32_t lb = a > s0 ? (uint32_t)(a - s0) : 0u;
            const uint32_t hb = x < s0 + 31 ? (uint32_t)(x - s0) : 31u;
            const uint32_t m = (uint32_t)(((2ull << hb) - 1) & ~((1ull << lb) - 1));
            keep &= ~m;
        }
    }
    return keep;
}

__device__ __forceinline__ void stage_tile(const DevGenome &g, int64_t w0, uint32_t *s_lo, uint32_t *s_hi,
                                           uint32_t *s_ot) {
    for (int i = threadIdx.x; i < kLdsWords; i += blockDim.x) {
        const int64_t w = w0 - kHaloL + i;
        const bool in = w >= 0 && (uint64_t)w < g.nwords;
        s_lo[i] = in ? g.lo[w] : 0u;
        s_hi[i] = in ? g.hi[w] : 0u;
        s_ot[i] = in ? g.ot[w] : ~0u;
    }
}

// exclusive scan of `v` over the 256-thread workgroup; *total gets the sum
__device__ __forceinline__ uint32_t block_scan(uint32_t v, uint32_t *s_wsum, uint32_t *total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
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

__device__ __forceinline__ void emit_hit(DevHit *hits, unsigned long long *count, uint64_t cap, uint64_t pos,
                                         uint32_t key, uint32_t mm, uint32_t strand, uint32_t qid) {
    const unsigned long long slot = atomicAdd(count, 1ull);
    if (slot < cap) {
        DevHit h;
        h.pos = pos;
        h.key = key;
        h.mm = (uint16_t)mm;
        h.strand = (uint8_t)strand;
        h.query = (uint8_t)qid;
        hits[slot] = h;
    }
}

// Window of NW words at key position cbase for one candidate, reverse
// complemented for strand 1, as class planes.  `bits` is the bit offset of
// the first word to read in the plane arrays p_*; rd(i) reads word i.
template <int NW, typename Rd>
__device__ __forceinline__ void window_classes(Rd rd, int64_t bit, uint32_t strand, uint32_t (&cT)[NW],
                                               uint32_t (&cC)[NW], uint32_t (&cA)[NW], uint32_t (&cG)[NW],
                                               uint32_t (&cO)[NW]) {
    uint32_t l[NW], h[NW], o[NW];
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
            h2[w] = __builtin_bitreverse32(h[NW - 1 - w]) ^ ~o2[w];  // complement: hi ^= 1 on bases
        }
#pragma unroll
        for (int w = 0; w < NW; w++) { l[w] = l2[w]; h[w] = h2[w] & ~o2[w]; o[w] = o2[w]; }
    }
#pragma unroll
    for (int w = 0; w < NW; w++) {
        cT[w] = ~(l[w] | h[w] | o[w]);
        cC[w] = l[w] & ~h[w];
        cA[w] = h[w] & ~l[w];
        cG[w] = l[w] & h[w];
        cO[w] = o[w];
    }
}

// Key tables are read-only for a launch: reading them through the constant
// address space lets the compiler use scalar loads (one per wavefront).
typedef const __attribute__((address_space(4))) uint32_t *kptr_t;

template <int NW, typename KP>
__device__ __forceinline__ uint32_t key_mismatches(KP kw, const uint32_t (&cT)[NW],
                                                   const uint32_t (&cC)[NW], const uint32_t (&cA)[NW],
                                                   const uint32_t (&cG)[NW], const uint32_t (&cO)[NW]) {
    uint32_t mm = 0;
#pragma unroll
    for (int w = 0; w < NW; w++) {
        const uint32_t x = (cT[w] & kw[kClsT * NW + w]) | (cC[w] & kw[kClsC * NW + w]) |
                           (cA[w] & kw[kClsA * NW + w]) | (cG[w] & kw[kClsG * NW + w]) |
                           (cO[w] & kw[kClsO * NW + w]);
        mm += __popc(x);
    }
    return mm;
}

}  // namespace

// ------------------------------------------------------------------ plain path

struct FinderArgs {
    DevGenome g;
    PatternDesc pd;
    uint64_t slab_lo, slab_hi;  // window starts [slab_lo, slab_hi), slab_lo tile aligned
    uint32_t *loci;             // locus - slab_lo
    uint8_t *flags;             // 0 both strands, 1 forward only, 2 reverse only
    unsigned long long *count;  // [0] loci written, [1] (locus, strand) candidates
    uint64_t cap;
};

__global__ __launch_bounds__(256) void finder_kernel(FinderArgs a) {
    __shared__ uint32_t s_lo[kLdsWords], s_hi[kLdsWords], s_ot[kLdsWords];
    __shared__ uint32_t s_wsum[4], s_wsum2[4];
    __shared__ unsigned long long s_base;
    const int t = threadIdx.x;
    const int64_t w0 = (int64_t)(a.slab_lo >> 5) + (int64_t)blockIdx.x * kTileWords;
    stage_tile(a.g, w0, s_lo, s_hi, s_ot);
    __syncthreads();
    const uint64_t t0 = (uint64_t)w0 * 32, t1 = t0 + kTilePos;
    const uint64_t s0 = t0 + 32ull * t;
    uint32_t fwd, rev;
    finder_masks(a.pd, s_lo, s_hi, s_ot, t + kHaloL, &fwd, &rev);
    const uint32_t keep = range_mask(s0, a.slab_lo, a.slab_hi) & poison_mask(a.g, a.pd.P, t0, t1, s0);
    fwd &= keep;
    rev &= keep;
    const uint32_t any = fwd | rev;
    uint32_t total, nstrand;
    const uint32_t off = block_scan(__popc(any), s_wsum, &total);
    (void)block_scan(__popc(fwd) + __popc(rev), s_wsum2, &nstrand);
    if (t == 0) {
        s_base = total ? atomicAdd(a.count, (unsigned long long)total) : 0ull;
        if (nstrand) atomicAdd(a.count + 1, (unsigned long long)nstrand);
    }
    __syncthreads();
    uint64_t o = s_base + off;
    uint32_t m = any;
    while (m) {
        const uint32_t b = __builtin_ctz(m);
        m &= m - 1;
        if (o < a.cap) {
            a.loci[o] = (uint32_t)(s0 + b - a.slab_lo);
            const uint32_t fb = (fwd >> b) & 1, rb = (rev >> b) & 1;
            a.flags[o] = (uint8_t)(fb && rb ? 0 : fb ? 1 : 2);
        }
        o++;
    }
}

struct ComparerArgs {
    DevGenome g;
    uint32_t P, cbase;
    uint32_t key[16];  // one compare key, key_stride(NW) dwords
    uint32_t key_index, qid;
    uint64_t slab_lo;
    const uint32_t *loci;
    const uint8_t *flags;
    uint64_t n;
    DevHit *hits;
    unsigned long long *count;
    uint64_t cap;
};

template <int NW>
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
    const uint32_t thr = a.key[5 * NW];
#pragma unroll
    for (uint32_t strand = 0; strand < 2; strand++) {
        if (strand == 0 ? !(fl == 0 || fl == 1) : !(fl == 0 || fl == 2)) continue;
        const int64_t bit = strand == 0 ? (int64_t)pos + a.cbase
                                        : (int64_t)pos + (int64_t)a.P - 32 * NW - (int64_t)a.cbase;
        uint32_t cT[NW], cC[NW], cA[NW], cG[NW], cO[NW];
        window_classes<NW>(rd, bit, strand, cT, cC, cA, cG, cO);
        const uint32_t mm = key_mismatches<NW>(a.key, cT, cC, cA, cG, cO);
        if (mm <= thr) emit_hit(a.hits, a.count, a.cap, pos, a.key_index, mm, strand, a.qid);
    }
}

// --------------------------------------------------------------- batched path

struct FusedArgs {
    DevGenome g;
    PatternDesc pd;
    uint32_t cbase, nkeys, qid, pad;
    const uint32_t *keys;  // nkeys * key_stride(NW) dwords
    uint64_t limit;        // valid window starts are [0, limit)
    DevHit *hits;
    unsigned long long *count;
    unsigned long long *cand_count;
    uint64_t cap;
};

template <int NW>
__global__ __launch_bounds__(256) void fused_kernel(FusedArgs a) {
    __shared__ uint32_t s_lo[kLdsWords], s_hi[kLdsWords], s_ot[kLdsWords];
    __shared__ uint16_t s_list[2 * kTilePos];
    __shared__ uint32_t s_wsum[4];
    const int t = threadIdx.x;
    const int64_t w0 = (int64_t)blockIdx.x * kTileWords;
    stage_tile(a.g, w0, s_lo, s_hi, s_ot);
    __syncthreads();
    const uint64_t t0 = (uint64_t)w0 * 32, t1 = t0 + kTilePos;
    const uint64_t s0 = t0 + 32ull * t;
    const uint32_t P = a.pd.P;
    uint32_t fwd, rev;
    finder_masks(a.pd, s_lo, s_hi, s_ot, t + kHaloL, &fwd, &rev);
    const uint32_t keep = range_mask(s0, 0, a.limit) & poison_mask(a.g, P, t0, t1, s0);
    fwd &= keep;
    rev &= keep;
    uint32_t n;
    uint32_t off = block_scan(__popc(fwd) + __popc(rev), s_wsum, &n);
    {
        uint32_t m = fwd;
        while (m) {
            const uint32_t b = __builtin_ctz(m);
            m &= m - 1;
            s_list[off++] = (uint16_t)((t * 32 + b) << 1);
        }
        m = rev;
        while (m) {
            const uint32_t b = __builtin_ctz(m);
            m &= m - 1;
            s_list[off++] = (uint16_t)(((t * 32 + b) << 1) | 1);
        }
    }
    if (t == 0 && n) atomicAdd(a.cand_count, (unsigned long long)n);
    __syncthreads();

    const int lane = t & 63, wave = t >> 6;
    auto rd = [&](int64_t wi, uint32_t &l, uint32_t &h, uint32_t &o) {
        l = s_lo[wi];
        h = s_hi[wi];
        o = s_ot[wi];
    };
    const uint32_t stride = key_stride(NW);
    for (uint32_t base = (uint32_t)wave * 64; base < n; base += 256) {
        const uint32_t idx = base + lane;
        if (idx < n) {
            const uint32_t e = s_list[idx];
            const uint32_t ls = e >> 1, strand = e & 1;
            const int64_t bit = strand == 0 ? (int64_t)ls + a.cbase + 32 * kHaloL
                                            : (int64_t)ls + P - 32 * NW - (int64_t)a.cbase + 32 * kHaloL;
            uint32_t cT[NW], cC[NW], cA[NW], cG[NW], cO[NW];
            window_classes<NW>(rd, bit, strand, cT, cC, cA, cG, cO);
            const uint64_t pos = t0 + ls;
            kptr_t kp = (kptr_t)a.keys;
#pragma unroll 4
            for (uint32_t k = 0; k < a.nkeys; k++, kp += stride) {
                const uint32_t mm = key_mismatches<NW>(kp, cT, cC, cA, cG, cO);
                if (mm <= kp[5 * NW]) emit_hit(a.hits, a.count, a.cap, pos, k, mm, strand, a.qid);
            }
        }
    }
}

// ------------------------------------------------------------------ launchers

hipError_t launch_finder(const DevGenome &g, const PatternDesc &pd, uint64_t slab_lo, uint64_t slab_hi,
                         uint32_t *loci, uint8_t *flags, unsigned long long *count, uint64_t cap,
                         hipStream_t s) {
    if (slab_hi <= slab_lo) return hipSuccess;
    FinderArgs a{g, pd, slab_lo, slab_hi, loci, flags, count, cap};
    const uint64_t tiles = (slab_hi - slab_lo + kTilePos - 1) / kTilePos;
    hipLaunchKernelGGL(finder_kernel, dim3((uint32_t)tiles), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_comparer(const DevGenome &g, uint32_t P, int nw, uint32_t cbase, const uint32_t *key,
                           uint32_t key_index, uint32_t qid, uint64_t slab_lo, const uint32_t *loci,
                           const uint8_t *flags, uint64_t n, DevHit *hits, unsigned long long *count, uint64_t cap,
                           hipStream_t s) {
    if (n == 0) return hipSuccess;
    ComparerArgs a{};
    a.g = g;
    a.P = P;
    a.cbase = cbase;
    for (int i = 0; i < key_stride(nw); i++) a.key[i] = key[i];
    a.key_index = key_index;
    a.qid = qid;
    a.slab_lo = slab_lo;
    a.loci = loci;
    a.flags = flags;
    a.n = n;
    a.hits = hits;
    a.count = count;
    a.cap = cap;
    const uint64_t blocks = (n + 255) / 256;
    if (nw == 1) hipLaunchKernelGGL(comparer_kernel<1>, dim3((uint32_t)blocks), dim3(256), 0, s, a);
    else hipLaunchKernelGGL(comparer_kernel<2>, dim3((uint32_t)blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_fused(const DevGenome &g, const PatternDesc &pd, int nw, uint32_t cbase, const uint32_t *keys,
                        uint32_t nkeys, uint32_t qid, DevHit *hits, unsigned long long *count,
                        unsigned long long *cand_count, uint64_t cap, hipStream_t s) {
    const uint64_t limit = g.total >= pd.P ? g.total - pd.P + 1 : 0;
    if (limit == 0) return hipSuccess;
    FusedArgs a{};
    a.g = g;
    a.pd = pd;
    a.cbase = cbase;
    a.nkeys = nkeys;
    a.qid = qid;
    a.keys = keys;
    a.limit = limit;
    a.hits = hits;
    a.count = count;
    a.cand_count = cand_count;
    a.cap = cap;
    const uint64_t tiles = (limit + kTilePos - 1) / kTilePos;
    if (nw == 1) hipLaunchKernelGGL(fused_kernel<1>, dim3((uint32_t)tiles), dim3(256), 0, s, a);
    else hipLaunchKernelGGL(fused_kernel<2>, dim3((uint32_t)tiles), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace cof
