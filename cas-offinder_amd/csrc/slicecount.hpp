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

// z = x + y, bit-sliced (six planes; the sum must stay below 64): a ripple
// of full adders.
COF_SLICE_FN void slice_add(const uint32_t (&x)[6], const uint32_t (&y)[6], uint32_t (&z)[6]) {
    uint32_t carry = 0;
#pragma unroll
    for (int w = 0; w < 6; w++) {
        const uint32_t a = x[w], b = y[w];
#if defined(__HIP_DEVICE_COMPILE__)
        z[w] = __builtin_amdgcn_bitop3_b32(a, b, carry, 0x96);
        carry = __builtin_amdgcn_bitop3_b32(a, b, carry, 0xE8);
#else
        z[w] = a ^ b ^ carry;
        carry = (a & b) | (carry & (a | b));
#endif
    }
}

// Lanes whose count c is at most t (both six bit planes): the count exceeds
// t at the highest plane where they differ with c's bit set.
COF_SLICE_FN uint32_t slice_le(const uint32_t (&c)[6], const uint32_t (&t)[6]) {
    uint32_t gt = 0, eq = ~0u;
#pragma unroll
    for (int w = 5; w >= 0; w--) {
        gt |= eq & c[w] & ~t[w];
        eq &= ~(c[w] ^ t[w]);
    }
    return ~gt;
}

}  // namespace cof
