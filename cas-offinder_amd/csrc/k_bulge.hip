// Batched bulge groups: bulge_kernel (device-side DNA/RNA-bulge expansion).
#include "kbulge.hpp"

namespace cof {

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
template <int NW, bool GEN, int NS, bool QUEUE, bool PAD>
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
    extern __shared__ __align__(16) uint32_t s_sep[];  // separator bitmap, kLdsWords words (PAD)
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
                const uint32_t room = pad_room<PAD>(a, ce, s_sep);
                uint32_t M[NS][NW];
#pragma unroll
                for (int u = 0; u < NS; u++)
#pragma unroll
                    for (int w = 0; w < NW; w++) M[u][w] = qm[lane * NS * NW + u * NW + w];
                const uint32_t vw = kp[KP];
                const kptr_t vb = tab + (vw & kVbOffsetMask);
                const TileSink sink = bulge_sink<PAD>(a, t0, ce, tab, vw, stage_ref(s_stage));
                const bool ok = !PAD || (vw >> kVbPadShift) <= room;
#pragma unroll
                for (int u = 0; u < NS; u++) {
                    if (ok && ((fam >> u) & 1)) {
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
        const uint32_t room = pad_room<PAD>(a, e[0], s_sep);
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
                        const uint32_t vw = kp[KP];
                        const kptr_t vb = tab + (vw & kVbOffsetMask);
                        const TileSink sink = bulge_sink<PAD>(a, t0, e[0], tab, vw, stage_ref(s_stage));
                        const bool ok = !PAD || (vw >> kVbPadShift) <= room;
#pragma unroll
                        for (int u = 0; u < NS; u++) {
                            if (ok && f[u] < 0) {
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
    }, 0u, nullptr, NoFin{}, [&] {
        if (PAD) stage_separators(a.g, (uint64_t)blockIdx.x * a.tile_stride, s_sep);
    });
}

template <int NW, bool GEN, int NS>
static void bulge_variant(int variant, dim3 grid, const BulgeArgs &a, hipStream_t s) {
    note_kernel("bulge_kernel", NW, GEN, NS, variant != 1, a.pad != 0);
    const uint32_t lds = a.pad ? kSepBytes : 0u;
    if (variant == 1 && a.pad) hipLaunchKernelGGL((bulge_kernel<NW, GEN, NS, false, true>), grid, dim3(kThreads), lds, s, a);
    else if (variant == 1) hipLaunchKernelGGL((bulge_kernel<NW, GEN, NS, false, false>), grid, dim3(kThreads), lds, s, a);
    else if (a.pad) hipLaunchKernelGGL((bulge_kernel<NW, GEN, NS, true, true>), grid, dim3(kThreads), lds, s, a);
    else hipLaunchKernelGGL((bulge_kernel<NW, GEN, NS, true, false>), grid, dim3(kThreads), lds, s, a);
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
                        const uint32_t *gap, bool rev, uint32_t pad, const uint32_t *table, uint32_t nunits,
                        uint64_t *hkey,
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
    a.pad = pad;
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

}  // namespace cof
