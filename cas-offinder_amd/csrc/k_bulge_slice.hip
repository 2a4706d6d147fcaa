// Batched bulge groups behind a bit-sliced filter of the guides:
// bulge_slice_kernel.
//
// bulge_kernel counts each (candidate, guide) pair in its own lane.  Here a
// lane holds one candidate and a 32-bit word holds 32 guides: per care
// position k of the guides (one contiguous run, kn <= 32) the word of the 32
// guides' mismatches against the candidate at window alignment u is
//   M_u[k] = (GL[k] ^ WL[k + d_u]) | (GH[k] ^ WH[k + d_u])
// with GL / GH the guides' code bit planes at k (one dword each, scalar
// loads) and WL / WH the candidate's code bits at that window position
// broadcast to all 32 bits.  A variant of the family of alignment u (DNA or
// RNA bulge of one size, any split j) compares guide positions below j at
// one alignment and those from j + gap at the other (cas-offinder.cpp:
// 650-682), so its mismatches are at least popc(M_0 & M_u) - gap; the family
// cannot hit unless the count of the planes M_0[k] & M_u[k] is at most
// threshold + gap, and the no-bulge key (popc(M_0) <= its threshold) is
// covered by the count of the guide's first family, which can only be
// smaller.  Slots are grouped by their family sets (a table may merge
// queries of different bulge sizes; a set may start a new word, leaving
// empty slots that never pass) and a word skips a family none of its
// guides has.  The
// counts of 32 guides come from one carry-save adder network over the kn
// planes (slicecount.hpp) and one bit-sliced comparison with the guides'
// bounds.  A family whose bounds in the word are large (slots are sorted
// by bound) is bounded on the two halves of its split range instead, like
// bulge_family's segments: far fewer pairs then reach the exact counts.
// A genome N reads as base T here (no non-ACGT plane): a count can
// only drop, so the filter stays a lower bound.  The (candidate, guide)
// pairs that pass go to a per-wave LDS queue and are counted exactly, one
// pair per lane, with bulge_kernel's family filters and every split of a
// passing family counted in one loop (bulge_family_flat: the lanes hold
// different guides, so segment bounds would diverge).
#include "kbulge.hpp"
#include "slicecount.hpp"

namespace cof {

struct SliceArgs {
    BulgeArgs b;
    SliceDesc sd;
    const uint32_t *slices;  // word blocks, then slot records (sd.slot_off)
};

// Per-wave queue of passing (candidate, guide) pairs: drained 64 at a time
// (one per lane) whenever it holds that many, the rest at the end of each
// candidate pass, so that drains inside the guide-word loop run full waves.
constexpr int kSliceQueue = 128;

// DN / DP: window shifts -1..-DN and +1..+DP of the filter families (the
// alignments of the table other than 0); KN: care positions of the guide
// words.  The round widths (16, 20, ..., 32) hold care runs of kn in
// (KN - 4, KN] and mask the positions past the run; the widths 21 and 23
// hold exactly their run (no mask).
// Three workgroups per CU where the registers allow it without spilling
// (care runs of up to 24 positions), two for 28 with one family.
template <int DN, int DP, int KN>
constexpr int slice_min_blocks() { return KN <= 24 ? 3 : DN + DP == 1 && KN <= 28 ? 2 : 1; }

template <int DN, int DP, int KN, bool PAD>
__global__ __launch_bounds__(kThreads, (slice_min_blocks<DN, DP, KN>())) void bulge_slice_kernel(SliceArgs sa) {
    constexpr bool CUT = KN % 4 == 0;  // a round width: the care run may be shorter
    constexpr int NF = DN + DP, NS = NF + 1, OFF = DN > DP ? DN : DP, PN = KN + 2 * OFF;
    constexpr int FW = PN <= 32 ? 1 : 2;  // frame words of the broadcast window positions
    __shared__ uint32_t s_lo[kLdsWords], s_hi[kLdsWords], s_ot[kLdsWords];
    __shared__ uint16_t s_list[kListCap];
    __shared__ uint32_t s_wsum[4];
    __shared__ HitStageBuf<kStageCap> s_stage;
    __shared__ uint32_t s_q[4][kSliceQueue];  // per wave: slot << 16 | candidate entry
    extern __shared__ __align__(16) uint32_t s_slot[];  // nslots x kBulgeSlotWords
    const BulgeArgs &a = sa.b;
    const SliceDesc &sd = sa.sd;
    uint32_t *s_sep = s_slot + sd.nslots * kBulgeSlotWords;  // separator bitmap, kLdsWords words (PAD)
    for (uint32_t i = threadIdx.x; i < sd.nslots * kBulgeSlotWords / 4; i += kThreads)
        reinterpret_cast<uint4 *>(s_slot)[i] = reinterpret_cast<const uint4 *>(sa.slices + sd.slot_off)[i];
    // visible after tile_candidates' first barrier
    constexpr int FB = bulge_family_words(1);
    const kptr_t tab = (kptr_t)a.table;
    const kptr_t blocks = (kptr_t)sa.slices;
    const uint4 *slot4 = reinterpret_cast<const uint4 *>(s_slot);
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t *q = s_q[wave];
    uint32_t qn = 0;  // queued pairs of this wave (uniform); the tile stays staged across its passes

    // exact counts of the first n queued pairs, one per lane
    auto drain = [&](uint64_t t0, auto rd, uint32_t n) __attribute__((always_inline)) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (lane < n) {
            const uint32_t item = q[lane];
            const uint32_t slot = item >> 16, ce = item & 0xFFFFu;
            const uint32_t room = pad_room<PAD>(a, ce, s_sep);
            uint32_t l[NS][1], h[NS][1], o[NS][1];
            bulge_windows<1, NS>(a, rd, ce, l, h, o);
            const uint4 s0 = slot4[2 * slot], s1 = slot4[2 * slot + 1];
            const uint32_t bias[5] = {s0.w, s1.x, s1.y, s1.z, s1.w};
            uint32_t M[NS];
            int32_t fl[NS];
            int32_t acc = 0;
#pragma unroll
            for (int u = 0; u < NS; u++) {
                const uint32_t lc = l[u][0] & sd.care, hc = h[u][0] & sd.care, oc = o[u][0] & sd.care;
                M[u] = __builtin_amdgcn_bitop3_b32(lc, s0.x, __builtin_amdgcn_bitop3_b32(hc, s0.y, oc, kXorOr),
                                                   kXorOr);
            }
#pragma unroll
            for (int u = 0; u < NS; u++) {
                fl[u] = (int32_t)(bias[u] + (uint32_t)__popc(u == 0 ? M[0] : M[u] & M[0]));
                acc |= fl[u];
            }
            if (acc < 0 && (!PAD || (s0.z >> kVbPadShift) <= room)) {
                const TileSink sink = bulge_sink<PAD>(a, t0, ce, tab, s0.z, stage_ref(s_stage));
                const kptr_t vb = tab + (s0.z & kVbOffsetMask);
#pragma unroll
                for (int u = 0; u < NS; u++) {
                    if (fl[u] < 0) {
                        const bool lft0 = u == 0 || a.rev;
                        const uint32_t left[1] = {lft0 ? M[0] : M[u]}, right[1] = {lft0 ? M[u] : M[0]};
                        bulge_family_flat<1>(tab, vb + 4 + FB * u, left, right, a.gap[u], bias[u], sink);
                    }
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
    };

    tile_candidates<1, true>(
        a.g, a.pd, a.limit, a.tile_stride, a.cand_count, s_lo, s_hi, s_ot, s_list, s_wsum, stage_ref(s_stage),
        a.hkey, a.hmm, a.count, a.cap,
        [&](uint64_t t0, auto rd, const uint32_t (&e)[1], const bool (&valid)[1]) {
            const uint32_t ce = e[0], ls = ce >> 1, strand = ce & 1;
            // frame: bit p = window position fbase + p (reverse complemented
            // for strand 1), broadcast per position
            const int64_t bit = strand == 0 ? (int64_t)ls + sd.fbase + 32 * kHaloL
                                            : (int64_t)ls + a.pd.P - 32 * FW - sd.fbase + 32 * kHaloL;
            uint32_t fl[FW], fh[FW], fo[FW];
            window_planes<FW>(rd, bit, strand, fl, fh, fo);
            uint32_t WL[PN], WH[PN];
#pragma unroll
            for (int p = 0; p < PN; p++) {
                WL[p] = (uint32_t)__builtin_amdgcn_sbfe((int32_t)fl[p / 32], p % 32, 1);
                WH[p] = (uint32_t)__builtin_amdgcn_sbfe((int32_t)fh[p / 32], p % 32, 1);
            }
            const uint32_t live = valid[0] ? ~0u : 0u;
            // masks of the last three care positions (all ones inside the run),
            // held in VGPRs: as uniform conditions they became spilled lane
            // masks read back with v_readlane in every word
            uint32_t keep[3];
#pragma unroll
            for (int j = 0; j < 3; j++) {
                keep[j] = (uint32_t)(KN - 3 + j) < sd.kn ? ~0u : 0u;
                if (CUT) asm volatile("" : "+v"(keep[j]));
            }
            for (uint32_t w = 0; w < sd.nwords; w++) {
                const kptr_t blk = blocks + w * sd.wstride;
                uint32_t M0[KN];
#pragma unroll
                for (int k = 0; k < KN; k++) {
                    M0[k] = __builtin_amdgcn_bitop3_b32(blk[k], WL[k + OFF], blk[KN + k] ^ WH[k + OFF], kXorOr);
                    if (CUT && k >= KN - 3) M0[k] &= keep[k - (KN - 3)];  // positions past the care run
                }
                uint32_t pass = 0;
                // bit f: family f bounded on the halves of its split range (uniform)
                const uint32_t fines = blk[2 * KN + (kSliceTPlanes + 1) * NF];
                uint32_t c0[6] = {};
                if (fines) {
                    if (a.rev) slice_half_shared<KN, true>(M0, c0);
                    else slice_half_shared<KN, false>(M0, c0);
                }
#pragma unroll
                for (int f = 0; f < NF; f++) {
                    const int d = f < DN ? -(f + 1) : f - DN + 1;
                    const kptr_t tp = blk + 2 * KN + (kSliceTPlanes + 1) * f;
                    if (tp[kSliceTPlanes] == 0) continue;  // no guide of the word has this family (uniform)
                    const uint32_t t[6] = {tp[0], tp[1], tp[2], tp[3], tp[4], tp[5]};
                    // guides' mismatches at alignment u, care position k
                    auto mu = [&](int k) {
                        return __builtin_amdgcn_bitop3_b32(blk[k], WL[k + OFF + d], blk[KN + k] ^ WH[k + OFF + d],
                                                           kXorOr);
                    };
                    uint32_t le;
                    if (!((fines >> f) & 1u)) {
                        // every split: mismatches at both alignments, at most bound + gap of them
                        uint32_t b[KN], c[6];
#pragma unroll
                        for (int k = 0; k < KN; k++) b[k] = M0[k] & mu(k);
                        slice_count<KN>(b, c);
                        le = slice_le(c, t);
                    } else {
                        // each half of the family's split range bounded on its own
                        // (slice_halves), M_0's count shared by the families in c0
                        uint32_t h1[6], h2[6];
                        if (a.rev) slice_halves<KN, true, CUT>(M0, mu, keep, c0, h1, h2);
                        else slice_halves<KN, false, CUT>(M0, mu, keep, c0, h1, h2);
                        le = slice_le(h1, t) | slice_le(h2, t);
                    }
                    pass |= le & tp[kSliceTPlanes];
                }
                pass &= live;
                // queue the passing (candidate, guide) pairs, one per lane per
                // round (qn < 64 before a round, so the queue holds its np <= 64)
                for (;;) {
                    const uint64_t m = __ballot(pass != 0);
                    if (!m) break;
                    const uint32_t np = (uint32_t)__popcll(m);
                    if (pass) {
                        const uint32_t g = (uint32_t)__builtin_ctz(pass);
                        pass &= pass - 1;
                        const uint32_t at =
                            qn + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
                        q[at] = ((32 * w + g) << 16) | ce;
                    }
                    qn += np;
                    if (qn >= 64) {
                        drain(t0, rd, 64);  // entries 0..63; the rest moves to the front
                        const uint32_t rest = qn - 64, v = lane < rest ? q[64 + lane] : 0u;
                        if (lane < rest) q[lane] = v;
                        qn = rest;
                    }
                }
            }
        },
        0u, nullptr,
        [&](uint64_t t0, auto rd) {
            if (qn) drain(t0, rd, qn);
            qn = 0;
        },
        [&] {
            if (PAD) stage_separators(a.g, (uint64_t)blockIdx.x * a.tile_stride, s_sep);
        });
}

template <int DN, int DP, int KN>
static hipError_t slice_launch(dim3 grid, const SliceArgs &a, uint32_t lds, hipStream_t s) {
    const bool pad = a.b.pad != 0;
    note_kernel("bulge_slice_kernel", DN, DP, KN, pad);
    if (pad) hipLaunchKernelGGL((bulge_slice_kernel<DN, DP, KN, true>), grid, dim3(kThreads), lds, s, a);
    else hipLaunchKernelGGL((bulge_slice_kernel<DN, DP, KN, false>), grid, dim3(kThreads), lds, s, a);
    return hipGetLastError();
}

template <int DN, int DP>
static hipError_t slice_kn(dim3 grid, const SliceArgs &a, uint32_t lds, hipStream_t s) {
    switch (a.sd.KN) {
        case 16: return slice_launch<DN, DP, 16>(grid, a, lds, s);
        case 20: return slice_launch<DN, DP, 20>(grid, a, lds, s);
        case 21: return slice_launch<DN, DP, 21>(grid, a, lds, s);
        case 23: return slice_launch<DN, DP, 23>(grid, a, lds, s);
        case 24: return slice_launch<DN, DP, 24>(grid, a, lds, s);
        case 28: return slice_launch<DN, DP, 28>(grid, a, lds, s);
        case 32: return slice_launch<DN, DP, 32>(grid, a, lds, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_bulge_slice(const DevGenome &g, const PatternDesc &pd, int ns, const int32_t *cb,
                              const uint32_t *gap, bool rev, uint32_t pad, const uint32_t *table, const SliceDesc &sd,
                              const uint32_t *slices, uint64_t *hkey, uint16_t *hmm, unsigned long long *count,
                              unsigned long long *cand_count, uint64_t cap, hipStream_t s, uint32_t tile_stride,
                              uint32_t max_tiles) {
    const uint64_t limit = g.total >= pd.P ? g.total - pd.P + 1 : 0;
    if (limit == 0 || sd.nslots == 0) return hipSuccess;
    // the kernel reads nwords blocks of wstride dwords and nslots slot records
    // after slot_off, 2 * KN + 7 * (dn + dp) dwords per block, and care
    // positions kn of the KN it is built for
    if (tile_stride == 0 || !slice_desc_valid(sd, ns) || (uintptr_t)slices % 16) return hipErrorInvalidValue;
    SliceArgs a{};
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
    a.sd = sd;
    a.slices = slices;
    const dim3 grid = tile_grid(limit, tile_stride, max_tiles);
    const uint32_t lds = sd.nslots * kBulgeSlotWords * 4 + (pad ? kSepBytes : 0u);
    switch (sd.dn * 3 + sd.dp) {
        case 1: return slice_kn<0, 1>(grid, a, lds, s);
        case 2: return slice_kn<0, 2>(grid, a, lds, s);
        case 3: return slice_kn<1, 0>(grid, a, lds, s);
        case 4: return slice_kn<1, 1>(grid, a, lds, s);
        case 6: return slice_kn<2, 0>(grid, a, lds, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace cof
