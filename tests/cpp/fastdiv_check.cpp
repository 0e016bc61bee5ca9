// Host check of the division-by-invariant-integer used for bucket indexes: fast_mod(n, d) must
// equal n % d for every tested n and d (exhaustive small d, edge n, random pairs).
#include <cstdio>
#include <cstdlib>
#include <random>

#include "dyc_device.hpp"

int main()
{
    using namespace dyc;
    std::mt19937_64 rng(12345);
    const uint32_t edge_n[] = {0u, 1u, 2u, 3u, 0x7fffffffu, 0x80000000u, 0xfffffffeu, 0xffffffffu, 294967290u};
    long bad = 0, checked = 0;
    auto check = [&](uint32_t n, uint32_t d) {
        FastDiv f = make_fastdiv(d);
        uint32_t r = fast_mod(n, f.d, f.m, f.s1, f.s2);
        ++checked;
        if (r != n % d) {
            if (bad < 10) printf("mismatch n=%u d=%u got %u want %u\n", n, d, r, n % d);
            ++bad;
        }
    };
    for (uint32_t d = 1; d < 5000; ++d) {
        for (uint32_t n : edge_n) check(n, d);
        for (int i = 0; i < 200; ++i) check((uint32_t)rng(), d);
    }
    const uint32_t big_d[] = {36424u, 72818u, 145636u, 4660218u, 9320436u, (1u << 29) + 11u, (1u << 30) - 1u,
                              (1u << 31) - 1u, 65536u, 1u << 30};
    for (uint32_t d : big_d) {
        for (uint32_t n : edge_n) check(n, d);
        for (int i = 0; i < 200000; ++i) check((uint32_t)rng(), d);
        for (uint32_t k = 0; k < 4; ++k) { check(d * k, d); if (d * k) check(d * k - 1, d); }
    }
    for (int i = 0; i < 2000000; ++i) {
        uint32_t d = (uint32_t)(rng() % 0x7fffffffu) + 1u;
        check((uint32_t)rng(), d);
    }
    printf("checked %ld bad %ld\n", checked, bad);
    return bad ? 1 : 0;
}
