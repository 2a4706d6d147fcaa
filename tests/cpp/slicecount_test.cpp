// Host check of the bit-sliced counting network (cas-offinder_amd/csrc/
// slicecount.hpp): for every plane count N in 1..32 and random planes, the
// six count planes equal each lane's popcount, slice_le equals the per-lane
// comparison with random bounds and slice_add the per-lane sum (with and
// without a carry-in plane).  The half bounds of a bulge family
// (slice_half_shared + slice_halves) equal their per-lane formula and are
// at most every split's mismatch count plus the gap, with random planes
// past the care run.  Prints "ok" or the first failure.
#include <cstdio>
#include <random>
#include <utility>

#include "../../cas-offinder_amd/csrc/slicecount.hpp"

using namespace cof;

template <int N>
static bool check(std::mt19937_64 &rng) {
    for (int rep = 0; rep < 2000; rep++) {
        uint32_t v[N], keep[N];
        const int density = rng() % 4;  // sparse to dense planes
        for (int i = 0; i < N; i++) {
            uint32_t x = (uint32_t)rng();
            for (int d = 0; d < density; d++) x |= (uint32_t)rng();
            if (rep % 7 == 0) x = ~0u;
            v[i] = keep[i] = x;
        }
        uint32_t c[6], t[6] = {};
        slice_count<N>(v, c);
        uint32_t tv[32];
        for (int g = 0; g < 32; g++) {
            tv[g] = (uint32_t)(rng() % 64);
            for (int w = 0; w < 6; w++) t[w] |= ((tv[g] >> w) & 1u) << g;
        }
        const uint32_t le = slice_le(c, t);
        uint32_t sum[6], sumc[6];
        slice_add(c, t, sum);
        const uint32_t cin = (uint32_t)rng();
        slice_add(c, t, sumc, cin);
        for (int g = 0; g < 32; g++) {
            int want = 0, got = 0;
            for (int i = 0; i < N; i++) want += (keep[i] >> g) & 1;
            for (int w = 0; w < 6; w++) got |= (int)((c[w] >> g) & 1) << w;
            if (got != want) {
                printf("count N=%d lane %d: %d vs %d\n", N, g, got, want);
                return false;
            }
            int s2 = 0;
            for (int w = 0; w < 6; w++) s2 |= (int)((sum[w] >> g) & 1) << w;
            if (want + (int)tv[g] < 64 && s2 != want + (int)tv[g]) {
                printf("add N=%d lane %d: %d + %u gave %d\n", N, g, want, tv[g], s2);
                return false;
            }
            int s3 = 0;
            for (int w = 0; w < 6; w++) s3 |= (int)((sumc[w] >> g) & 1) << w;
            const int wc = want + (int)tv[g] + (int)((cin >> g) & 1);
            if (wc < 64 && s3 != wc) {
                printf("add+cin N=%d lane %d: gave %d, want %d\n", N, g, s3, wc);
                return false;
            }
            if ((int)((le >> g) & 1) != (want <= (int)tv[g])) {
                printf("le N=%d lane %d: count %d bound %u\n", N, g, want, tv[g]);
                return false;
            }
        }
    }
    return true;
}

template <int... N>
static bool all(std::mt19937_64 &rng, std::integer_sequence<int, N...>) {
    return (check<N + 1>(rng) && ...);
}


// per-lane count of planes p over care positions [lo, hi)
template <int KN>
static int lane_count(const uint32_t (&p)[KN], int g, int lo, int hi) {
    int n = 0;
    for (int k = lo; k < hi; k++) n += (p[k] >> g) & 1;
    return n;
}

template <int KN, bool RV, bool CUT>
static bool check_halves(std::mt19937_64 &rng) {
    constexpr int S = KN / 2;
    for (int rep = 0; rep < 3000; rep++) {
        const int kn = CUT ? KN - (int)(rng() % 4) : KN;  // care run of KN - 3 .. KN positions (CUT)
        uint32_t m0[KN], mu[KN];
        for (int k = 0; k < KN; k++) {
            // mismatch planes about as dense as a random window's (3 in 4), some sparse
            uint32_t x = (uint32_t)rng() | (uint32_t)rng(), y = (uint32_t)rng() | (uint32_t)rng();
            if (rep % 3 == 0) x &= (uint32_t)rng() & (uint32_t)rng();
            if (rep % 5 == 0) y &= (uint32_t)rng() & (uint32_t)rng();
            m0[k] = k < kn ? x : 0u;  // M_0 is zero past the care run
            mu[k] = y;                // M_u holds whatever the window has there
        }
        uint32_t c0[6], h1[6], h2[6];
        slice_half_shared<KN, RV>(m0, c0);
        uint32_t keep[3];
        for (int j = 0; j < 3; j++) keep[j] = KN - 3 + j < kn ? ~0u : 0u;
        slice_halves<KN, RV, CUT>(m0, [&](int k) { return mu[k]; }, keep, c0, h1, h2);
        uint32_t mu_in[KN], b[KN];
        for (int k = 0; k < KN; k++) {
            mu_in[k] = k < kn ? mu[k] : 0u;
            b[k] = m0[k] & mu[k];
        }
        const uint32_t(&left)[KN] = RV ? m0 : mu_in;
        const uint32_t(&right)[KN] = RV ? mu_in : m0;
        for (int g = 0; g < 32; g++) {
            int v1 = 0, v2 = 0;
            for (int w = 0; w < 6; w++) {
                v1 |= (int)((h1[w] >> g) & 1) << w;
                v2 |= (int)((h2[w] >> g) & 1) << w;
            }
            const int f1 = lane_count<KN>(b, g, 0, S + 1) + lane_count<KN>(right, g, S + 1, kn);
            const int f2 = lane_count<KN>(left, g, 0, S) + lane_count<KN>(b, g, S, kn);
            if (v1 != f1 || v2 != f2) {
                printf("halves KN=%d RV=%d CUT=%d kn=%d lane %d: (%d, %d) vs formula (%d, %d)\n", KN, RV, CUT, kn, g,
                       v1, v2, f1, f2);
                return false;
            }
            // every split j of a family of gap 0..2: mm(j) + gap bounds the half holding j
            for (int gap = 0; gap <= 2; gap++)
                for (int j = 0; j + gap <= kn; j++) {
                    const int mm = lane_count<KN>(left, g, 0, j) + lane_count<KN>(right, g, j + gap, kn);
                    const int h = j < S ? v1 : v2;
                    if (h > mm + gap) {
                        printf("halves KN=%d RV=%d kn=%d lane %d gap %d split %d: bound %d > %d + gap\n", KN, RV, kn,
                               g, gap, j, h, mm);
                        return false;
                    }
                }
        }
    }
    return true;
}

template <int KN, bool CUT>
static bool check_halves_both(std::mt19937_64 &rng) {
    return check_halves<KN, false, CUT>(rng) && check_halves<KN, true, CUT>(rng);
}

int main() {
    std::mt19937_64 rng(12345);
    if (!all(rng, std::make_integer_sequence<int, 32>{})) return 1;
    // the kernel's widths: masked round widths, the exact widths 21 and 23
    if (!(check_halves_both<16, true>(rng) && check_halves_both<20, true>(rng) && check_halves_both<24, true>(rng) &&
          check_halves_both<28, true>(rng) && check_halves_both<32, true>(rng) && check_halves_both<21, false>(rng) &&
          check_halves_both<23, false>(rng)))
        return 1;
    // op count of the networks the kernel uses: at most one full adder per input
    for (int n : {16, 20, 24, 28, 32}) {
        const CsaPlan p = csa_plan(n);
        if (p.nops > n) {
            printf("plan %d: %d ops\n", n, p.nops);
            return 1;
        }
    }
    printf("ok\n");
    return 0;
}
