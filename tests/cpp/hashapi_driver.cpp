// The reference's gtest cases (GPU/dyhash/Main_backup_0910.cpp, GPU/warp/Main.cpp), replayed
// through the hashAPI drop-in of include/dyc_hashapi.hpp. Exact checks replace the reference's
// "<= 2% mismatches" tolerance. Prints one line per case; exit status = number of failed cases.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <random>
#include <set>
#include <vector>

#include "dyc_hashapi.hpp"

using TYPE = hashAPI::TYPE;
static int failures = 0;

static void report(const char *name, bool ok)
{
    printf("%s %s\n", ok ? "PASS" : "FAIL", name);
    fflush(stdout);
    if (!ok) ++failures;
}

static std::vector<TYPE> unique_keys(size_t n, uint64_t seed)
{
    std::mt19937_64 rng(seed);
    std::set<TYPE> seen;
    std::vector<TYPE> out;
    while (out.size() < n) {
        TYPE k = (TYPE)(rng() % 0x7fffffffu) + 1u;  // 31-bit keys as in random_numbers.cpp:46
        if (seen.insert(k).second) out.push_back(k);
    }
    return out;
}

// HastTest.SizeTEST (Main_backup_0910.cpp:33-95, 208-217): values 3i+4, table sized 1.1 n
static bool size_test(TYPE n)
{
    auto key = unique_keys(n, 1000 + n);
    std::vector<TYPE> value(n), chck(n, 0);
    for (TYPE i = 0; i < n; ++i) value[i] = 3 * i + 3 + 1;
    hashAPI h((TYPE)(n * 1.1));
    h.hash_insert(key.data(), value.data(), n);
    h.hash_search(key.data(), chck.data(), n);
    return chck == value;
}

// HashBoundaryTest.InsertsameKeyValue (Main_backup_0910.cpp:236-275): 100 x (11, 9)
static bool same_key_test()
{
    const TYPE N = 100;
    std::vector<TYPE> key(N, 11), value(N, 9), search(N), chck(N, 0);
    std::vector<uint8_t> hit(N);
    for (TYPE i = 0; i < N; ++i) search[i] = i;
    hashAPI h(N);
    try {
        h.hash_insert(key.data(), value.data(), N);
    } catch (const dyc_error &e) {
        // one key has only 2 buckets = 32 slots: copies beyond that cannot be placed
        if (e.code() != DYC_ERR_PENDING) throw;
    }
    h.hash_search_hits(search.data(), chck.data(), hit.data(), N);
    bool ok = true;
    for (TYPE i = 0; i < N; ++i) ok &= (i == 11) ? (chck[i] == 9 && hit[i]) : (chck[i] == 0 && !hit[i]);
    std::fill(chck.begin(), chck.end(), 0);
    h.hash_search(key.data(), chck.data(), N);
    for (TYPE i = 0; i < N; ++i) ok &= chck[i] == 9;
    return ok;
}

// HashBoundaryTest.RandTestKV (Main_backup_0910.cpp:278-313) with unique keys
static bool rand_kv_test()
{
    const TYPE N = 100;
    auto key = unique_keys(N, 77);
    std::vector<TYPE> value(N), chck(N, 0);
    for (TYPE i = 0; i < N; ++i) value[i] = i;
    hashAPI h(N);
    h.hash_insert(key.data(), value.data(), N);
    h.hash_search(key.data(), chck.data(), N);
    return chck == value;
}

// HastTestRemove.Test1 (warp/Main.cpp:126-222): delete the first half, exact check
static bool remove_test(TYPE n)
{
    auto key = unique_keys(n, 500 + n);
    std::vector<TYPE> value(n), chck(n, 0), ans(n, 0);
    for (TYPE i = 0; i < n; ++i) value[i] = 3 * i + 3 + 1;
    hashAPI h((TYPE)(n * 1.25));
    h.hash_insert(key.data(), value.data(), n);
    h.hash_delete(key.data(), ans.data(), n / 2);
    bool ok = std::all_of(ans.begin(), ans.begin() + n / 2, [](TYPE a) { return a == 1; });
    h.hash_search(key.data(), chck.data(), n);
    for (TYPE i = 0; i < n; ++i) ok &= (i < n / 2) ? chck[i] == 0 : chck[i] == value[i];
    return ok;
}

// HashResizeTest.RandTestWithUpDown (Main_backup_0910.cpp:321-411): 1M keys, table 2M,
// resize_low twice with a find after each; plus resize_up (warp/Main.cpp:328-520)
static bool resize_test(TYPE n)
{
    auto key = unique_keys(n, 99);
    std::vector<TYPE> value(n), chck(n, 0);
    for (TYPE i = 0; i < n; ++i) value[i] = 3 * i + 3 + 1;
    hashAPI h(n * 2);
    h.hash_insert(key.data(), value.data(), n);
    h.hash_search(key.data(), chck.data(), n);
    bool ok = chck == value;
    for (int r = 0; r < 2; ++r) {
        h.resize_low();
        std::fill(chck.begin(), chck.end(), 0);
        h.hash_search(key.data(), chck.data(), n);
        ok &= chck == value;
    }
    h.resize_up();
    std::fill(chck.begin(), chck.end(), 0);
    h.hash_search(key.data(), chck.data(), n);
    ok &= chck == value;
    dyc_stats s = h.stats();
    ok &= s.count == n && s.pending == 0 && s.lost == 0 && s.resizes_down == 2 && s.resizes_up == 1;
    return ok;
}

// HashnewApiTest.test1 (warp/Main.cpp:530-611): batch insert + search, then batch delete
static bool batch_test()
{
    const TYPE batches = 8, n = batches * hashAPI::BATCH_SIZE;
    auto key = unique_keys(n, 4242);
    std::vector<TYPE> value(n), chck(n, 0);
    for (TYPE i = 0; i < n; ++i) value[i] = i + 1;
    hashAPI h(n * 2);
    h.set_data_to_gpu(key.data(), value.data(), n);
    for (TYPE b = 0; b < batches; ++b) {
        h.hash_insert_batch();
        h.hash_search_batch();
    }
    h.hash_search(key.data(), chck.data(), n);
    bool ok = chck == value;
    for (TYPE b = 0; b < batches / 2; ++b) h.hash_delete_batch();
    std::fill(chck.begin(), chck.end(), 0);
    h.hash_search(key.data(), chck.data(), n);
    for (TYPE i = 0; i < n; ++i) ok &= (i < n / 2) ? chck[i] == 0 : chck[i] == value[i];
    return ok;
}

// HashStreamTEST.Four (GPU/old_version/cuckoo_gtest/Main.cpp:37-138): one host thread, four
// streams, 1M keys per stream copied and inserted on its own stream, all in flight at once, into a
// table made for 5M slots; then every key is searched.
static bool stream_test(TYPE size, TYPE init_size, int nstreams)
{
    const TYPE n = size * (TYPE)nstreams;
    auto key = unique_keys(n, 4444);
    std::vector<TYPE> value(n), chck(n, 0);
    for (TYPE i = 0; i < n; ++i) value[i] = i + 1;
    hashAPI h(init_size);
    std::vector<hipStream_t> st(nstreams);
    std::vector<TYPE *> dk(nstreams), dv(nstreams);
    bool ok = true;
    for (int i = 0; i < nstreams; ++i) {
        ok &= hipStreamCreate(&st[i]) == hipSuccess;
        ok &= hipMalloc(&dk[i], sizeof(TYPE) * size) == hipSuccess && hipMalloc(&dv[i], sizeof(TYPE) * size) == hipSuccess;
    }
    if (!ok) return false;
    for (int i = 0; i < nstreams; ++i) {
        ok &= hipMemcpyAsync(dk[i], key.data() + (size_t)i * size, sizeof(TYPE) * size, hipMemcpyHostToDevice, st[i]) ==
              hipSuccess;
        ok &= hipMemcpyAsync(dv[i], value.data() + (size_t)i * size, sizeof(TYPE) * size, hipMemcpyHostToDevice,
                             st[i]) == hipSuccess;
        h.hash_insert_async(dk[i], dv[i], size, st[i]);
    }
    h.sync();
    h.hash_search(key.data(), chck.data(), n);
    ok &= chck == value;
    const dyc_stats s = h.stats();
    ok &= s.count == n && s.pending == 0 && s.lost == 0;
    for (int i = 0; i < nstreams; ++i) {
        (void)hipFree(dk[i]);
        (void)hipFree(dv[i]);
        (void)hipStreamDestroy(st[i]);
    }
    return ok;
}

// The batch API's resize policy (api.cpp:353-369, 466-506) on a table sized far below its data:
// 24 batches in, then 20 batches out, with the reference's fill bounds 0.8 / 0.5 (test_uilt.h:1-2)
// and auto_resize off. Prints the subtable sizes and the count after every batch, which the
// caller compares with the oracle's replay of the same batch sizes, and checks every key.
static bool batch_resize_trace()
{
    const TYPE batches = 24, out = 20, n = batches * hashAPI::BATCH_SIZE;
    auto key = unique_keys(n, 777);
    std::vector<TYPE> value(n), chck(n, 0);
    for (TYPE i = 0; i < n; ++i) value[i] = i + 1;
    dyc_config cfg;
    dyc_config_init(&cfg);
    cfg.upper_bound = 0.8;
    cfg.lower_bound = 0.5;
    cfg.min_capacity = 0;
    hashAPI h(1u << 18, &cfg);
    h.set_data_to_gpu(key.data(), value.data(), n);
    auto show = [&](const char *op, TYPE b) {
        const dyc_stats s = h.stats();
        printf("SIZES %s %u", op, b);
        for (uint32_t i = 0; i < s.subtables; ++i) printf(" %u", s.buckets[i]);
        printf(" count %llu\n", (unsigned long long)s.count);
    };
    show("start", 0);
    for (TYPE b = 0; b < batches; ++b) {
        h.hash_insert_batch();
        show("insert", b);
    }
    h.hash_search(key.data(), chck.data(), n);
    bool ok = chck == value;
    for (TYPE b = 0; b < out; ++b) {
        h.hash_delete_batch();
        show("delete", b);
    }
    std::fill(chck.begin(), chck.end(), 0);
    h.hash_search(key.data(), chck.data(), n);
    const TYPE gone = out * hashAPI::BATCH_SIZE;
    for (TYPE i = 0; i < n; ++i) ok &= (i < gone) ? chck[i] == 0 : chck[i] == value[i];
    const dyc_stats s = h.stats();
    return ok && s.count == n - gone && s.pending == 0 && s.lost == 0 && s.resizes_up > 0 && s.resizes_down > 0;
}

int main(int argc, char **argv)
{
    if (argc > 1 && strcmp(argv[1], "batch-resize") == 0) {
        report("BatchApiResize", batch_resize_trace());
        printf("%d failed\n", failures);
        return failures;
    }
    const TYPE sizes[] = {100, 200, 1024, 2048, 4096, 1u << 16, 1u << 20};  // Main_backup_0910.cpp:210
    for (TYPE n : sizes) {
        char name[64];
        snprintf(name, sizeof name, "SizeTEST/%u", n);
        report(name, size_test(n));
    }
    report("InsertsameKeyValue", same_key_test());
    report("RandTestKV", rand_kv_test());
    const TYPE rsizes[] = {100, 4096, 1u << 16};
    for (TYPE n : rsizes) {
        char name[64];
        snprintf(name, sizeof name, "Remove/%u", n);
        report(name, remove_test(n));
    }
    report("RandTestWithUpDown", resize_test(1000000));
    report("BatchApi", batch_test());
    report("HashStreamTEST.Four", stream_test(1000000, 5000000, 4));
    printf("%d failed\n", failures);
    return failures;
}
