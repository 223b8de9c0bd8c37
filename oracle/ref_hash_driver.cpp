// Driver for the reference's own integer-hash RNG (src/pscbfm/rngs/Hash.h,
// included from where it lies under the reference tree by oracle/ref.mk).
// Reads "seed subsequence" pairs from stdin and prints the first rng32() of a
// fresh Rngs::Hash set up the way the check kernel does it (.cu:958-964).
// Test infrastructure only: its output pins the oracle's orc_int_hash.
#include <cstdio>

#include "Hash.h"

int main()
{
    unsigned long long seed = 0, sub = 0;
    while (std::scanf("%llu %llu", &seed, &sub) == 2) {
        Rngs::Hash rng;
        rng.setSubsequence(sub);
        rng.setSeed(seed);
        std::printf("%u\n", rng.rng32());
    }
    return 0;
}
