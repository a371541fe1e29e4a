// Test infrastructure: prints rocRAND's philox4x32_10 output for counters read
// from stdin ("c0 c1 c2 c3 k0 k1" in hex, one per line).  Host-only; touches no GPU.
// Used to pin the oracle's Philox against AMD's open-source engine.
#include <cstdio>
#include <rocrand/rocrand_philox4x32_10.h>

int main() {
    unsigned c0, c1, c2, c3, k0, k1;
    while (std::scanf("%x %x %x %x %x %x", &c0, &c1, &c2, &c3, &k0, &k1) == 6) {
        unsigned long long key = (unsigned long long)k0 | ((unsigned long long)k1 << 32);
        unsigned long long subseq = (unsigned long long)c2 | ((unsigned long long)c3 << 32);
        unsigned long long offset = 4ull * ((unsigned long long)c0 | ((unsigned long long)c1 << 32));
        rocrand_device::philox4x32_10_engine eng(key, subseq, offset);
        uint4 r = eng.next4();
        std::printf("%08x %08x %08x %08x\n", r.x, r.y, r.z, r.w);
    }
    return 0;
}
