/* Exercises every oracle entry point on small tables, for the sanitizer builds (ASan + UBSan,
 * TSan): sequential and OpenMP inserts and finds (the multithreaded CPU baseline shares the
 * image through 64-bit CAS), erase, upsize and downsize with residual re-insertion, and the
 * wide-key table. Exits 0 and prints "oracle stress ok" when every check holds. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "dyc_oracle.h"

#define CHECK(c)                                                      \
    do {                                                              \
        if (!(c)) {                                                   \
            fprintf(stderr, "check failed line %d: %s\n", __LINE__, #c); \
            exit(1);                                                  \
        }                                                             \
    } while (0)

static uint64_t rng_state = 0x9E3779B97F4A7C15ull;
static uint64_t next64(void)
{
    rng_state ^= rng_state << 13;
    rng_state ^= rng_state >> 7;
    rng_state ^= rng_state << 17;
    return rng_state;
}

static int cmp_u32(const void *a, const void *b)
{
    const uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    return x < y ? -1 : x > y;
}

/* n distinct non-reserved keys in shuffled order */
static uint32_t *unique_keys(uint64_t n)
{
    uint64_t m = n + n / 4 + 16;
    uint32_t *k = malloc(m * sizeof *k);
    for (;;) {
        for (uint64_t i = 0; i < m; ++i) {
            uint32_t x;
            do x = (uint32_t)next64(); while (x == 0u || x == 0xFFFFFFFFu);
            k[i] = x;
        }
        qsort(k, m, sizeof *k, cmp_u32);
        uint64_t u = 0;
        for (uint64_t i = 0; i < m; ++i)
            if (u == 0 || k[i] != k[u - 1]) k[u++] = k[i];
        if (u >= n) break;
    }
    for (uint64_t i = n - 1; i > 0; --i) {
        const uint64_t j = next64() % (i + 1);
        const uint32_t t = k[i];
        k[i] = k[j];
        k[j] = t;
    }
    return k;
}

static void table_alloc(orc_table *tb, uint32_t d, uint32_t nb)
{
    memset(tb, 0, sizeof *tb);
    tb->d = d;
    for (uint32_t t = 0; t < d; ++t) {
        tb->nb[t] = nb;
        tb->sub[t] = calloc((size_t)nb * 16, sizeof(uint64_t));
    }
}

static void table_free(orc_table *tb)
{
    for (uint32_t t = 0; t < tb->d; ++t) free(tb->sub[t]);
}

static void check_all_found(const orc_table *tb, const uint32_t *k, const uint32_t *v, uint64_t n, int threads)
{
    uint32_t *got = malloc(n * sizeof *got);
    uint8_t *hit = malloc(n);
    if (threads > 1)
        orc_find_mt(tb, k, n, got, hit, threads);
    else
        orc_find(tb, k, n, got, hit, NULL);
    for (uint64_t i = 0; i < n; ++i) CHECK(hit[i] == 1 && got[i] == v[i]);
    free(got);
    free(hit);
}

int main(void)
{
    const uint64_t n = 60000;
    uint32_t *k = unique_keys(n);
    uint32_t *v = malloc(n * sizeof *v);
    for (uint64_t i = 0; i < n; ++i) v[i] = (uint32_t)(k[i] * 2654435761u + 7u);
    uint64_t *fail = malloc(n * sizeof *fail);
    const uint32_t d = 4, nb = orc_initial_buckets_d((uint64_t)(n / 0.85), d);

    /* sequential and multithreaded inserts of the same keys: same key->value set */
    orc_table a, b;
    table_alloc(&a, d, nb);
    table_alloc(&b, d, nb);
    uint64_t rej = 0, kicks = 0;
    CHECK(orc_insert(&a, k, v, n, 15, 128, NULL, fail, n, &rej) == 0 && rej == 0);
    CHECK(orc_insert_mt(&b, k, v, n, 15, 128, fail, n, &kicks, &rej, 8) == 0 && rej == 0);
    CHECK(orc_check_invariant(&a) == 0 && orc_check_invariant(&b) == 0);
    CHECK(orc_count(&a) == n && orc_count(&b) == n);
    check_all_found(&a, k, v, n, 1);
    check_all_found(&b, k, v, n, 8);

    /* erase the first half; the rest stays */
    uint8_t *del = malloc(n);
    CHECK(orc_erase(&a, k, n / 2, del) == n / 2);
    for (uint64_t i = 0; i < n / 2; ++i) CHECK(del[i] == 1);
    check_all_found(&a, k + n / 2, v + n / 2, n - n / 2, 1);

    /* upsize every subtable of b, then downsize each back with residual re-insertion */
    for (uint32_t t = 0; t < d; ++t) {
        uint64_t *big = calloc((size_t)b.nb[t] * 32, sizeof(uint64_t));
        orc_upsize(b.sub[t], b.nb[t], big, t);
        free(b.sub[t]);
        b.sub[t] = big;
        b.nb[t] *= 2;
        CHECK(orc_check_invariant(&b) == 0);
    }
    check_all_found(&b, k, v, n, 1);
    for (uint32_t t = 0; t < d; ++t) {
        const uint32_t half = b.nb[t] / 2;
        uint64_t *small = calloc((size_t)half * 16, sizeof(uint64_t));
        uint64_t *res = malloc((size_t)half * 16 * sizeof(uint64_t));
        const uint64_t nres = orc_downsize_compact(b.sub[t], b.nb[t], small, res, (uint64_t)half * 16);
        free(b.sub[t]);
        b.sub[t] = small;
        b.nb[t] = half;
        for (uint64_t i = 0; i < nres; ++i) {
            uint32_t kk = 0;
            uint64_t carry = 0;
            CHECK(orc_insert_entry(&b, res[i], t, 15, 128, &kk, &carry) == 1);
        }
        free(res);
        CHECK(orc_check_invariant(&b) == 0);
    }
    check_all_found(&b, k, v, n, 8);
    CHECK(orc_count(&b) == n);

    /* wide keys */
    orc_wide w;
    memset(&w, 0, sizeof w);
    w.d = d;
    const uint64_t wn = 20000;
    const uint32_t wnb = orc_initial_buckets_d((uint64_t)(wn / 0.85), d);
    for (uint32_t t = 0; t < d; ++t) {
        w.nb[t] = wnb;
        w.keys[t] = calloc((size_t)wnb * 16, sizeof(uint64_t));
        w.vals[t] = calloc((size_t)wnb * 16, sizeof(uint64_t));
    }
    uint64_t *wk = malloc(wn * sizeof *wk), *wv = malloc(wn * sizeof *wv), *fk = malloc(wn * 8), *fv = malloc(wn * 8);
    for (uint64_t i = 0; i < wn; ++i) {
        wk[i] = ((uint64_t)k[i] << 32) | k[n - 1 - i];
        wv[i] = ~wk[i];
    }
    CHECK(orc_wide_insert(&w, wk, wv, wn, 15, 128, NULL, fk, fv, wn, NULL) == 0);
    CHECK(orc_wide_check_invariant(&w) == 0);
    uint64_t *wgot = malloc(wn * 8);
    uint8_t *whit = malloc(wn);
    orc_wide_find(&w, wk, wn, wgot, whit);
    for (uint64_t i = 0; i < wn; ++i) CHECK(whit[i] == 1 && wgot[i] == wv[i]);
    CHECK(orc_wide_erase(&w, wk, wn / 3, whit) == wn / 3);
    for (uint32_t t = 0; t < d; ++t) {
        free(w.keys[t]);
        free(w.vals[t]);
    }
    free(wk), free(wv), free(fk), free(fv), free(wgot), free(whit);

    table_free(&a);
    table_free(&b);
    free(k), free(v), free(fail), free(del);
    printf("oracle stress ok\n");
    return 0;
}
