// Host-only cross-check of the oracle's Philox4x32-10 against rocRAND's
// philox4x32_10_engine (header-only, compiled for the host).  Prints, for each
// (seed, subsequence) pair, the first 8 words of the engine's stream:
// the words of Philox(counter=(c,0,sub_lo,sub_hi), key=(seed_lo,seed_hi)) for c=0,1.
#include <cstdio>
#include <rocrand/rocrand_philox4x32_10.h>
int main() {
    const unsigned long long seeds[] = {0ull, 1ull, 42ull, 0x123456789abcdefull};
    const unsigned long long subs[] = {0ull, 3ull, 0x100000002ull};
    for (unsigned long long s : seeds)
        for (unsigned long long q : subs) {
            rocrand_device::philox4x32_10_engine eng(s, q, 0);
            std::printf("%llu %llu", s, q);
            for (int i = 0; i < 8; ++i) std::printf(" %u", eng.next());
            std::printf("\n");
        }
    return 0;
}
