// Host-only harness of the library's worker threads (dyc_threads.hpp), for the sanitizer builds:
// the deferred-free helper (ReclaimerT, here over malloc/free) and the copy pool of the
// host-array calls. Several owners run at once, as several tables do. Exits 0 and prints
// "threads ok" when every block was freed exactly once and every copy is exact.
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <atomic>
#include <thread>
#include <vector>

#include "dyc_threads.hpp"

namespace {

std::atomic<uint64_t> g_freed{0};
std::atomic<uint64_t> g_bytes{0};

struct CountingFree {
    void operator()(int device, void *p) const
    {
        // the block carries its own size and the device it was deferred with
        const uint64_t *w = static_cast<const uint64_t *>(p);
        if ((int)w[1] != device) abort();
        g_bytes += w[0];
        ++g_freed;
        free(p);
    }
};

void reclaim_owner(int id, uint64_t blocks, uint64_t *deferred)
{
    dyc::host::ReclaimerT<CountingFree> r;
    uint64_t x = 0x9E3779B97F4A7C15ull * (uint64_t)(id + 1);
    for (uint64_t i = 0; i < blocks; ++i) {
        x ^= x << 13, x ^= x >> 7, x ^= x << 17;
        const uint64_t bytes = 16 + x % 4096;
        uint64_t *p = static_cast<uint64_t *>(malloc(bytes));
        p[0] = bytes;
        p[1] = (uint64_t)id;
        r.defer(id, p);
        *deferred += bytes;
        if (i % 997 == 996) r.drain();  // a table's destroy, then a new life of the same helper
    }
}  // the destructor drains the rest

bool copy_owner(unsigned workers, int id)
{
    dyc::host::CopyPool pool(workers);
    std::vector<unsigned char> src(24u << 20), dst(24u << 20);
    for (size_t i = 0; i < src.size(); ++i) src[i] = (unsigned char)(i * 131u + (unsigned)id);
    const size_t sizes[] = {0, 1, 63, 4096, (4u << 20) - 1, 4u << 20, (4u << 20) + 1, 9999999, 24u << 20};
    for (int rep = 0; rep < 3; ++rep)
        for (size_t n : sizes) {
            const size_t off = (size_t)rep * 7;
            if (n + off > src.size()) continue;
            memset(dst.data(), 0xEE, dst.size());
            pool.copy(dst.data() + off, src.data() + off, n);
            if (memcmp(dst.data() + off, src.data() + off, n) != 0) return false;
            if (off && dst[off - 1] != 0xEE) return false;
            if (n + off < dst.size() && dst[n + off] != 0xEE) return false;
        }
    return true;
}

}  // namespace

int main()
{
    const int owners = 4;
    const uint64_t blocks = 20000;
    std::vector<uint64_t> deferred(owners, 0);
    std::vector<std::thread> th;
    for (int i = 0; i < owners; ++i) th.emplace_back(reclaim_owner, i, blocks, &deferred[i]);
    std::atomic<int> copy_bad{0};
    for (int i = 0; i < 2; ++i)
        th.emplace_back([i, &copy_bad] {
            if (!copy_owner(i == 0 ? 7u : 0u, i)) ++copy_bad;
        });
    for (auto &t : th) t.join();
    uint64_t total = 0;
    for (uint64_t d : deferred) total += d;
    if (g_freed != owners * blocks || g_bytes != total || copy_bad) {
        printf("freed %llu of %llu, bytes %llu of %llu, bad copies %d\n", (unsigned long long)g_freed.load(),
               (unsigned long long)(owners * blocks), (unsigned long long)g_bytes.load(), (unsigned long long)total,
               copy_bad.load());
        return 1;
    }
    printf("threads ok\n");
    return 0;
}
