// Host build of the device gain_bucket helpers: the loop-free version must
// equal the literal loop (jet_refiner.hpp:326-356) on every input checked.
// Prints the number of inputs compared; exits 1 on the first difference.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>

#include "jet_types.h"

static long long checked = 0;

static void check(int64_t g, int64_t w) {
    const int32_t a = jet::gain_bucket(g, w), b = jet::gain_bucket_fast(g, w);
    checked++;
    if (a != b) {
        std::printf("MISMATCH g=%lld w=%lld loop=%d fast=%d\n", (long long)g, (long long)w, a, b);
        std::exit(1);
    }
}

int main() {
    // dense small integers: every weight up to 64, gains in [-40000, 40000]
    for (int64_t w = 1; w <= 64; w++)
        for (int64_t g = -40000; g <= 40000; g++) check(g, w);
    // each threshold's neighbourhood: with w = 2^s the quotient g / w is
    // exact, so g = t * w + d walks the floats next to threshold t once w
    // makes t * w an integer
    for (int s = 0; s <= 40; s++) {
        const int64_t w = (int64_t)1 << s;
        for (float t : jet::GB_UP)
            for (int64_t d = -3; d <= 3; d++) check(-((int64_t)std::llround((double)t * (double)w) + d), w);
        for (float t : jet::GB_DOWN)
            for (int64_t d = -3; d <= 3; d++) check(-((int64_t)std::llround((double)t * (double)w) + d), w);
    }
    // random 64-bit gains and weights, wide and narrow
    std::mt19937_64 rs(7);
    for (int i = 0; i < 2000000; i++) {
        const int64_t w = 1 + (int64_t)(rs() % (i % 2 ? 1000000007ull : 97ull));
        const int64_t span = (i % 3 == 0) ? (1ll << 40) : (i % 3 == 1 ? 100000 : 200);
        const int64_t g = (int64_t)(rs() % (uint64_t)(2 * span + 1)) - span;
        check(g, w);
    }
    std::printf("%lld\n", checked);
    return 0;
}
