// Host check of the bit-sliced counting network (cas-offinder_amd/csrc/
// slicecount.hpp): for every plane count N in 1..32 and random planes, the
// six count planes equal each lane's popcount, slice_le equals the per-lane
// comparison with random bounds and slice_add the per-lane sum.  Prints "ok" or the first failure.
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
        uint32_t sum[6];
        slice_add(c, t, sum);
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

int main() {
    std::mt19937_64 rng(12345);
    if (!all(rng, std::make_integer_sequence<int, 32>{})) return 1;
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
