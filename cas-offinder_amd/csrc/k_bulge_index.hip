// Batched bulge groups behind the bit-set seed index of the guides: bulge_index_kernel.
#include "kbulge.hpp"

namespace cof {

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
// then per slot [glo, ghi, variant block word, family filter bias per
// alignment] (kBulgeSlotWords words).
struct BulgeIndexArgs {
    BulgeArgs b;
    MaskDesc md;
    const uint32_t *idx;
};

// Register budget: six waves per SIMD for one-word rows (they fit 80
// VGPRs without spilling), five for two, three for four.
template <int NK, int NS, int NP, bool PAD>
__global__ __launch_bounds__(kThreads, NK == 4 ? 3 : NK == 2 ? 5 : 6) void bulge_index_kernel(BulgeIndexArgs ia) {
    __shared__ uint32_t s_lo[kLdsWords], s_hi[kLdsWords], s_ot[kLdsWords];
    __shared__ uint16_t s_list[kListCap];
    __shared__ uint32_t s_wsum[4];
    __shared__ HitStageBuf<kStageCap> s_stage;
    extern __shared__ __align__(16) uint32_t s_idx[];
    const BulgeArgs &a = ia.b;
    uint32_t *s_sep = s_idx + ia.md.nwords;  // separator bitmap, kLdsWords words (PAD)
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
        const uint32_t room = pad_room<PAD>(a, ce, s_sep);
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
                if (acc >= 0 || (PAD && (s0.z >> kVbPadShift) > room)) continue;
                const kptr_t vb = tab + (s0.z & kVbOffsetMask);
                const TileSink sink = bulge_sink<PAD>(a, t0, ce, tab, s0.z, stage_ref(s_stage));
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
    }, 0u, nullptr, NoFin{}, [&] {
        if (PAD) stage_separators(a.g, (uint64_t)blockIdx.x * a.tile_stride, s_sep);
    });
}

template <int NK, int NS>
static hipError_t bulge_index_parts(uint32_t parts, dim3 grid, const BulgeIndexArgs &a, uint32_t lds, hipStream_t s) {
    const bool pad = a.b.pad != 0;
    note_kernel("bulge_index_kernel", NK, NS, parts <= 4 ? 4 : parts <= 6 ? 6 : 8, pad);
    if (parts <= 4) {
        if (pad) hipLaunchKernelGGL((bulge_index_kernel<NK, NS, 4, true>), grid, dim3(kThreads), lds, s, a);
        else hipLaunchKernelGGL((bulge_index_kernel<NK, NS, 4, false>), grid, dim3(kThreads), lds, s, a);
    } else if (parts <= 6) {
        if (pad) hipLaunchKernelGGL((bulge_index_kernel<NK, NS, 6, true>), grid, dim3(kThreads), lds, s, a);
        else hipLaunchKernelGGL((bulge_index_kernel<NK, NS, 6, false>), grid, dim3(kThreads), lds, s, a);
    } else {
        if (pad) hipLaunchKernelGGL((bulge_index_kernel<NK, NS, 8, true>), grid, dim3(kThreads), lds, s, a);
        else hipLaunchKernelGGL((bulge_index_kernel<NK, NS, 8, false>), grid, dim3(kThreads), lds, s, a);
    }
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
                              const uint32_t *gap, bool rev, uint32_t pad, const uint32_t *table, const MaskDesc &md,
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
    b.pad = pad;
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
    const uint32_t lds = md.nwords * 4 + (pad ? kSepBytes : 0u);
    switch (md.nk) {
        case 1: return bulge_index_ns<1>(ns, md.parts, grid, a, lds, s);
        case 2: return bulge_index_ns<2>(ns, md.parts, grid, a, lds, s);
        default: return bulge_index_ns<4>(ns, md.parts, grid, a, lds, s);
    }
}

}  // namespace cof
