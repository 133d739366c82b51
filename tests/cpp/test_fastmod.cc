// gt_fastmod / gt_fastmod30 (goetia_amd/csrc/gt_hash.h) against the % operator:
// table sizes of the tests and workloads, primes near 2^20..2^30, small moduli,
// and hashes at the edges of the 64-bit range and from a seeded generator.
// Built for the host with ROCm's clang and run by tests/test_fastmod.py.
#include <cstdio>
#include <cstdint>
#include <initializer_list>

#include "gt_hash.h"

static uint64_t next(uint64_t& x) {
    x += 0x9E3779B97F4A7C15ULL;
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

int main() {
    const uint64_t ps[] = {1, 2, 3, 7, 999983, 9999991, 1048573, 268435399, 268435337, 536870909,
                           1073741789, 1073741723, (1ULL << 30) - 1, 1ULL << 29};
    const uint64_t edges[] = {0, 1, 2, 0xFFFFFFFFULL, 0x100000000ULL, 0x7FFFFFFFFFFFFFFFULL, 0x8000000000000000ULL,
                              0xFFFFFFFFFFFFFFFEULL, 0xFFFFFFFFFFFFFFFFULL};
    uint64_t x = 12345, bad = 0, n = 0;
    for (uint64_t p : ps) {
        const uint64_t m = gt_mod_magic(p);
        for (uint64_t e : edges) {
            for (uint64_t d = 0; d < 3 * p && d < 4096; ++d) {
                for (uint64_t h : {e + d, e - d, (e / p) * p + d, (e / p) * p - d}) {
                    bad += gt_fastmod(h, p, m) != h % p;
                    bad += gt_fastmod30(h, (uint32_t)p, m) != h % p;
                    n += 2;
                }
            }
        }
        for (int i = 0; i < 2000000; ++i) {
            const uint64_t h = next(x);
            bad += gt_fastmod(h, p, m) != h % p;
            bad += gt_fastmod30(h, (uint32_t)p, m) != h % p;
            n += 2;
        }
    }
    std::printf("checked %llu, mismatches %llu\n", (unsigned long long)n, (unsigned long long)bad);
    return bad != 0;
}
