// Random alloc/release of the serving arena against a byte-ownership map.
#include <cstdio>
#include <random>
#include <vector>

#include "dyc_arena.hpp"

int main()
{
    using dyc::Arena;
    const uint64_t size = 1 << 20;  // 4096 units of 256 bytes
    Arena a;
    a.reset(size + 100);  // rounded down to the alignment
    if (a.size() != size) return printf("size %llu\n", (unsigned long long)a.size()), 1;
    std::vector<int> owner(size / Arena::kAlign, -1);
    std::vector<std::pair<uint64_t, uint64_t>> live;  // (offset, units)
    std::mt19937_64 rng(7);
    uint64_t used = 0;
    int id = 0, fails = 0;
    for (int step = 0; step < 200000; ++step) {
        if (live.empty() || rng() % 100 < 55) {
            const uint64_t bytes = 1 + rng() % (size / 16);
            const uint64_t off = a.alloc(bytes);
            const uint64_t units = (bytes + Arena::kAlign - 1) / Arena::kAlign;
            if (off == Arena::kNone) {
                // refused only if no free run of that many units exists
                uint64_t run = 0, best = 0;
                for (int o : owner) {
                    run = o < 0 ? run + 1 : 0;
                    best = run > best ? run : best;
                }
                if (best >= units) return printf("refused %llu units with a free run of %llu\n",
                                                 (unsigned long long)units, (unsigned long long)best), 1;
                ++fails;
                continue;
            }
            if (off % Arena::kAlign) return printf("misaligned %llu\n", (unsigned long long)off), 1;
            for (uint64_t u = off / Arena::kAlign; u < off / Arena::kAlign + units; ++u) {
                if (u >= owner.size() || owner[u] >= 0) return printf("overlap at unit %llu\n", (unsigned long long)u), 1;
                owner[u] = id;
            }
            ++id;
            live.emplace_back(off, units);
            used += units * Arena::kAlign;
        } else {
            const size_t k = rng() % live.size();
            const auto [off, units] = live[k];
            if (!a.release(off)) return printf("release refused\n"), 1;
            if (a.release(off)) return printf("double release accepted\n"), 1;
            for (uint64_t u = off / Arena::kAlign; u < off / Arena::kAlign + units; ++u) owner[u] = -1;
            live[k] = live.back();
            live.pop_back();
            used -= units * Arena::kAlign;
        }
        if (a.used() != used || a.free_bytes() != size - used) return printf("accounting off\n"), 1;
    }
    for (const auto &l : live) a.release(l.first);
    if (a.used() != 0 || a.free_blocks() != 1 || a.largest_free() != size) return printf("not coalesced\n"), 1;
    printf("arena ok: %d allocations, %d refused, peak %llu\n", id, fails, (unsigned long long)a.peak());
    return 0;
}
