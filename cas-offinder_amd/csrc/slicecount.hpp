// Bit-sliced counting for the bulge slice filter: N one-bit planes (bit g of
// plane k = input k of lane g) summed into six count planes (bit g of plane
// w = bit w of lane g's count) with a carry-save adder network, and the
// per-lane comparison of two bit-sliced numbers.  The network is a
// compile-time schedule of full and half adders (the number of 3:2 and 2:2
// compressions of column-wise popcount reduction), run by template
// recursion so that every plane index is a compile-time constant and the
// planes stay in registers.  Host-compilable, so that host tests check the
// same schedule against a plain popcount.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define COF_SLICE_FN __host__ __device__ __forceinline__
#else
#define COF_SLICE_FN inline
#endif

namespace cof {

struct CsaOp {
    int8_t kind, a, b, c;  // 3: full adder of planes a, b, c (sum in a, carry in b); 2: half adder of a, b
};
struct CsaPlan {
    int nops = 0;
    CsaOp ops[64] = {};
    int8_t out[6] = {-1, -1, -1, -1, -1, -1};  // plane holding count bit w (-1: always 0)
};

// Column compression: while a weight holds three planes, a full adder turns
// them into one plane of the same weight and a carry of the next; two left
// take a half adder.  Input planes are reused for sums and carries.
constexpr CsaPlan csa_plan(int n) {
    CsaPlan p{};
    int8_t col[8][64] = {};
    int cnt[8] = {};
    for (int i = 0; i < n; i++) col[0][cnt[0]++] = (int8_t)i;
    for (int w = 0; w < 7; w++) {
        while (cnt[w] >= 3) {
            const int8_t a = col[w][--cnt[w]], b = col[w][--cnt[w]], c = col[w][--cnt[w]];
            p.ops[p.nops++] = CsaOp{3, a, b, c};
            col[w][cnt[w]++] = a;
            col[w + 1][cnt[w + 1]++] = b;
        }
        if (cnt[w] == 2) {
            const int8_t a = col[w][--cnt[w]], b = col[w][--cnt[w]];
            p.ops[p.nops++] = CsaOp{2, a, b, 0};
            col[w][cnt[w]++] = a;
            col[w + 1][cnt[w + 1]++] = b;
        }
        if (w < 6) p.out[w] = cnt[w] ? col[w][0] : (int8_t)-1;
    }
    return p;
}

template <int N, int I>
COF_SLICE_FN void csa_run(uint32_t (&v)[N]) {
    constexpr CsaPlan P = csa_plan(N);
    if constexpr (I < P.nops) {
        constexpr CsaOp o = P.ops[I];
        const uint32_t a = v[o.a], b = v[o.b];
        if constexpr (o.kind == 3) {
            const uint32_t c = v[o.c];
#if defined(__HIP_DEVICE_COMPILE__)
            // one v_bitop3 each (truth tables over a*4 + b*2 + c): sum = XOR3, carry = majority
            v[o.a] = __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
            v[o.b] = __builtin_amdgcn_bitop3_b32(a, b, c, 0xE8);
#else
            v[o.a] = a ^ b ^ c;
            v[o.b] = (a & b) | (c & (a | b));
#endif
        } else {
            v[o.a] = a ^ b;
            v[o.b] = a & b;
        }
        csa_run<N, I + 1>(v);
    }
}

// c[w] = bit w of the per-lane count of the N planes of v (v is clobbered).
template <int N>
COF_SLICE_FN void slice_count(uint32_t (&v)[N], uint32_t (&c)[6]) {
    static_assert(N >= 1 && N <= 63, "at most 63 planes");
    csa_run<N, 0>(v);
    constexpr CsaPlan P = csa_plan(N);
#pragma unroll
    for (int w = 0; w < 6; w++) c[w] = P.out[w] >= 0 ? v[P.out[w] >= 0 ? P.out[w] : 0] : 0u;
}

// Count planes that can be set for a count of at most n: the bit width of n.
constexpr int slice_bits(int n) { return n <= 0 ? 0 : 1 + slice_bits(n >> 1); }

// z = x + y + cin, bit-sliced (six planes, cin one plane; the sum must stay
// below 64): a ripple of full adders over the B low planes, where both x
// and y are zero from plane B on (the default: all six).  The adders stop
// there: the carry is plane B and the planes above it are zero (the
// compiler does not fold v_bitop3 with zero operands, so adders over known
// zero planes would be issued).
template <int B = 6>
COF_SLICE_FN void slice_add(const uint32_t (&x)[6], const uint32_t (&y)[6], uint32_t (&z)[6], uint32_t cin = 0) {
    static_assert(B >= 1 && B <= 6, "one to six planes");
    uint32_t carry = cin;
#pragma unroll
    for (int w = 0; w < B; w++) {
        const uint32_t a = x[w], b = y[w];
#if defined(__HIP_DEVICE_COMPILE__)
        z[w] = __builtin_amdgcn_bitop3_b32(a, b, carry, 0x96);
        carry = __builtin_amdgcn_bitop3_b32(a, b, carry, 0xE8);
#else
        z[w] = a ^ b ^ carry;
        carry = (a & b) | (carry & (a | b));
#endif
    }
#pragma unroll
    for (int w = B; w < 6; w++) z[w] = w == B ? carry : 0u;
}

// Lanes whose count c is at most t (both six bit planes): c <= t exactly
// when c + (63 - t) carries nothing out of bit 5, so one carry chain of
// majorities of c, ~t and the carry (one v_bitop3 per bit) decides it.
COF_SLICE_FN uint32_t slice_le(const uint32_t (&c)[6], const uint32_t (&t)[6]) {
    uint32_t carry = c[0] & ~t[0];
#pragma unroll
    for (int w = 1; w < 6; w++) {
#if defined(__HIP_DEVICE_COMPILE__)
        carry = __builtin_amdgcn_bitop3_b32(c[w], t[w], carry, 0xB2);  // majority of c, ~t, carry
#else
        carry = (c[w] & ~t[w]) | (carry & (c[w] | ~t[w]));
#endif
    }
    return ~carry;
}

// ------------------------------------------------ half bounds of a bulge family
//
// A variant of a bulge family at split j counts the left part's mismatches
// at care positions below j and the right part's from j + gap (gap <= 2;
// cas-offinder.cpp:650-682), so with b = M_0 & M_u (a mismatch at both
// alignments) every position counts at least b there, except the gap
// positions (the filter bound adds gap to the threshold).  The splits of
// the family are bounded on two halves, S = KN / 2:
//   splits j < S: positions up to S hold b, those past S only the right part
//     (j + gap <= S + 1): h1 = count(b, k <= S) + count(right, k > S);
//   splits j >= S: positions below S only the left part:
//     h2 = count(left, k < S) + count(b, k >= S).
// The left part is M_u and the right part M_0 for a forward PAM (RV false),
// the other way round for a reversed one, so one of the two counts of M_0
// alone is the same for every family of a word: slice_half_shared computes
// it once (forward: M_0 over (S, KN); reversed: M_0 over [0, S)).
// Positions from kn on (KN - 4 < kn <= KN) lie past the guide's care run and
// count nothing: m0 is zero there, and mu(k) is masked where it enters a
// count alone (CUT: kn may be below KN, keep[j] is all ones where position
// KN - 3 + j is inside the run, else zero; without CUT kn = KN, as for the
// kernel's widths 21 and 23).
template <int KN, bool RV>
COF_SLICE_FN void slice_half_shared(const uint32_t (&m0)[KN], uint32_t (&c0)[6]) {
    constexpr int S = KN / 2;
    if constexpr (RV) {
        uint32_t v[S];
#pragma unroll
        for (int k = 0; k < S; k++) v[k] = m0[k];
        slice_count<S>(v, c0);
    } else {
        uint32_t v[KN - S - 1];
#pragma unroll
        for (int k = S + 1; k < KN; k++) v[k - S - 1] = m0[k];
        slice_count<KN - S - 1>(v, c0);
    }
}

// h1 / h2: count planes of the two half bounds of the family whose mismatch
// plane at care position k is mu(k); c0 from slice_half_shared.
template <int KN, bool RV, bool CUT, typename Mu>
COF_SLICE_FN void slice_halves(const uint32_t (&m0)[KN], Mu mu, const uint32_t (&keep)[3], const uint32_t (&c0)[6],
                               uint32_t (&h1)[6], uint32_t (&h2)[6]) {
    constexpr int S = KN / 2;
    uint32_t pa[S + 1], a1[6];  // b at k <= S
    uint32_t x1[6] = {};        // forward: left = M_u below S
    {
        uint32_t px[S];
#pragma unroll
        for (int k = 0; k < S; k++) {
            const uint32_t m = mu(k);
            pa[k] = m0[k] & m;
            px[k] = m;
        }
        if constexpr (!RV) slice_count<S>(px, x1);
    }
    const uint32_t bs = m0[S] & mu(S);
    pa[S] = bs;
    slice_count<S + 1>(pa, a1);
    uint32_t pe[KN - S - 1], e1[6];  // b at k > S
    uint32_t y1[6] = {};             // reversed: right = M_u past S
    {
        uint32_t py[KN - S - 1];
#pragma unroll
        for (int k = S + 1; k < KN; k++) {
            uint32_t m = mu(k);
            pe[k - S - 1] = m0[k] & m;
            if (CUT && k >= KN - 3) m &= keep[k - (KN - 3)];  // past the care run (kn > KN - 4)
            py[k - S - 1] = m;
        }
        if constexpr (RV) slice_count<KN - S - 1>(py, y1);
    }
    slice_count<KN - S - 1>(pe, e1);
    // every partial count is of at most S + 1 planes: its planes from
    // slice_bits(S + 1) on are zero
    constexpr int B = slice_bits(S + 1);
    if constexpr (RV) {
        slice_add<B>(a1, y1, h1);
        slice_add<B>(c0, e1, h2, bs);
    } else {
        slice_add<B>(a1, c0, h1);
        slice_add<B>(x1, e1, h2, bs);
    }
}

}  // namespace cof
