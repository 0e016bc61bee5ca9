// Host check of nth_set_bit16 (the bulk fill's slot choice): for every 16-bit mask and every rank
// below its popcount, the position equals the one found by clearing the lowest set bit r times.
#include <cstdio>

#include "dyc_device.hpp"

int main()
{
    long bad = 0, checked = 0;
    for (uint32_t m = 0; m < 65536u; ++m) {
        uint32_t x = m;
        for (uint32_t r = 0; x; ++r, x &= x - 1u) {
            const uint32_t want = (uint32_t)__builtin_ctz(x);
            const uint32_t got = dyc::nth_set_bit16(m, r);
            ++checked;
            if (got != want) {
                if (bad < 10) printf("mismatch m=%#x r=%u got %u want %u\n", m, r, got, want);
                ++bad;
            }
        }
    }
    printf("checked %ld bad %ld\n", checked, bad);
    return bad != 0;
}
