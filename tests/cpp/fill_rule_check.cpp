// Host print of the bulk fill's first-round rule (Desc::fr from desc_init) for every subtable
// count: "d fr" per line, for the test to check against its own search over the oracle's pairs.
#include <cstdio>

#include "dyc_device.hpp"

int main()
{
    for (uint32_t d = (uint32_t)dyc::kMinSubtables; d <= (uint32_t)dyc::kMaxSubtables; ++d) {
        dyc::Desc D;
        if (!dyc::desc_init(D, d)) return 1;
        printf("%u %u\n", d, D.fr);
    }
    return 0;
}
