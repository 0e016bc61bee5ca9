// dyc_device.hpp -- DyCuckoo table descriptor and per-key arithmetic shared by every kernel.
//
// Layout (unchanged from the reference, cuckoo.h:13-16,30-34): 4 subtables, each an array of
// 128-byte buckets of 16 uint64 slots; slot = (uint64)key << 32 | value, empty = 0.
// Keys 0 (empty marker) and 0xFFFFFFFF are reserved.
//
// All functions are __host__ __device__ so the host code computes the same numbers the kernels
// do (dump checks, the fast-division parameters).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dyc {

constexpr int kSubtables = 4;                  // dy_hash.h:32 TABLE_NUM
constexpr int kSlots = 16;                     // dy_hash.h:34 ELEM_NUM_P_gpu = 4
constexpr int kBucketBytes = kSlots * 8;       // 128
constexpr uint32_t kPrime = 294967291u;        // dy_hash.h:24 PRIME_uint
constexpr uint32_t kNoTable = 100u;            // cuckoo.cu:336 initial pre_table_no
constexpr uint32_t kDefaultKickBound = 15u;    // dy_hash.h:33 MAX_ITER

__host__ __device__ inline uint32_t key_of(uint64_t e) { return (uint32_t)(e >> 32); }
__host__ __device__ inline uint32_t val_of(uint64_t e) { return (uint32_t)e; }
__host__ __device__ inline uint64_t pack(uint32_t k, uint32_t v) { return ((uint64_t)k << 32) | v; }
__host__ __device__ inline bool reserved_key(uint32_t k) { return k == 0u || k == 0xffffffffu; }

// ---------------------------------------------------------------------------------- hashes
// The five integer hashes of cuckoo.cu:27-64,110-120 (Wang shift, Jenkins 6-shift, the affine
// h3, Wang shift-multiply, Jenkins 7-shift). The formulas are the reference's: they define
// bucket placement, so any change would break table-image parity.

__host__ __device__ inline uint32_t h1(uint32_t k)
{
    k = ~k + (k << 15);
    k ^= k >> 12;
    k += k << 2;
    k ^= k >> 4;
    k *= 2057u;
    k ^= k >> 16;
    return k;
}

__host__ __device__ inline uint32_t h2(uint32_t a)
{
    a = (a + 0x7ed55d16u) + (a << 12);
    a = (a ^ 0xc761c23cu) ^ (a >> 19);
    a = (a + 0x165667b1u) + (a << 5);
    a = (a + 0xd3a2646cu) ^ (a << 9);
    a = (a + 0xfd7046c5u) + (a << 3);
    a = (a ^ 0xb55a4f09u) ^ (a >> 16);
    return a;
}

__host__ __device__ inline uint32_t h3(uint32_t k) { return ((k ^ 59064253u) + 72355969u) % kPrime; }

__host__ __device__ inline uint32_t h4(uint32_t a)
{
    a = (a ^ 61u) ^ (a >> 16);
    a += a << 3;
    a ^= a >> 4;
    a *= 0x27d4eb2du;
    a ^= a >> 15;
    return a;
}

__host__ __device__ inline uint32_t h5(uint32_t a)
{
    a -= a << 6;
    a ^= a >> 17;
    a -= a << 9;
    a ^= a << 4;
    a -= a << 3;
    a ^= a << 10;
    a ^= a >> 15;
    return a;
}

// Subtable t's hash reduced mod PRIME_uint (cuckoo.cu:80-89). The four hashes are evaluated
// branch-free and selected: t differs lane to lane, and a divergent switch would run all four
// paths under masks anyway.
__host__ __device__ inline uint32_t hash_t(uint32_t k, uint32_t t)
{
    uint32_t a = h1(k), b = h2(k), c = h3(k), d = h4(k);
    uint32_t x = t == 0 ? a : (t == 1 ? b : (t == 2 ? c : d));
    return x % kPrime;
}

// ------------------------------------------------------------------------- pair selection
// Two-layer scheme (cuckoo.cu:190-213,310): a key may live only in the two subtables of pair
// h5(k) % 6; pairs 0..5 = (1,2) (0,1) (2,3) (0,2) (1,3) (0,3).
__host__ __device__ inline uint32_t pair_of(uint32_t k) { return h5(k) % 6u; }
__host__ __device__ inline uint32_t pos1(uint32_t p) { return (p & 1u) ? 0u : ((p & 2u) ? 2u : 1u); }
__host__ __device__ inline uint32_t pos2(uint32_t p)
{
    return (p & 1u) ? (p >> 1) + 1u : ((p & 2u) ? 3u : (p >> 2) + 2u);
}
__host__ __device__ inline uint32_t other_pos(uint32_t p, uint32_t pre) { return pre == pos1(p) ? pos2(p) : pos1(p); }

// Victim lane of the rescue phase that follows the reference kick bound: the top 4 bits of a
// multiplicative hash of the in-hand key and the iteration. (The reference drops the entry at
// its kick bound, cuckoo.cu:422-431.)
__host__ __device__ inline uint32_t rescue_lane(uint32_t k, uint32_t it)
{
    return (k * 0x9E3779B1u + it * 0x632BE5ABu) >> 28;
}

// ------------------------------------------------------------------ division by a runtime d
// n % d for the per-subtable bucket count d without a hardware divide: Granlund & Montgomery,
// "Division by Invariant Integers using Multiplication" (PLDI 1994), Fig. 4.1 (round-up
// multiplier with the add-and-shift fix-up), exact for every 32-bit n and 1 <= d < 2^31.
struct FastDiv {
    uint32_t d, m, s1, s2;
};

inline FastDiv make_fastdiv(uint32_t d)
{
    uint32_t l = 0;
    while (l < 32 && ((uint64_t)1 << l) < d) ++l;  // l = ceil(log2 d)
    FastDiv f;
    f.d = d;
    f.m = (uint32_t)((((uint64_t)1 << 32) * (((uint64_t)1 << l) - d)) / d + 1);
    f.s1 = l < 1 ? l : 1;
    f.s2 = l > 0 ? l - 1 : 0;
    return f;
}

__host__ __device__ inline uint32_t mulhi32(uint32_t a, uint32_t b)
{
#ifdef __HIP_DEVICE_COMPILE__
    return __umulhi(a, b);
#else
    return (uint32_t)(((uint64_t)a * b) >> 32);
#endif
}

__host__ __device__ inline uint32_t fast_div(uint32_t n, uint32_t m, uint32_t s1, uint32_t s2)
{
    uint32_t t1 = mulhi32(m, n);
    return (t1 + ((n - t1) >> s1)) >> s2;
}

__host__ __device__ inline uint32_t fast_mod(uint32_t n, uint32_t d, uint32_t m, uint32_t s1, uint32_t s2)
{
    return n - fast_div(n, m, s1, s2) * d;
}

// ------------------------------------------------------------------------ table descriptor
// Passed by value as a kernel argument (the reference keeps one process-global __constant__
// descriptor, cuckoo.cu:16, which limits a process to one table). Fields are uniform, so they
// live in SGPRs; per-lane selection by subtable is a v_cndmask chain.
struct Desc {
    uint64_t *sub[kSubtables];   // bucket arrays, nb[t] * 16 slots each
    uint32_t nb[kSubtables];     // buckets per subtable
    uint32_t m[kSubtables];      // FastDiv multipliers for nb[t]
    uint32_t s1[kSubtables];
    uint32_t s2[kSubtables];
};

template <typename T>
__host__ __device__ inline T sel4(uint32_t t, T a, T b, T c, T d)
{
    return t == 0 ? a : (t == 1 ? b : (t == 2 ? c : d));
}

__host__ __device__ inline uint32_t bucket_of(const Desc &D, uint32_t k, uint32_t t)
{
    uint32_t h = hash_t(k, t);
    return fast_mod(h, sel4(t, D.nb[0], D.nb[1], D.nb[2], D.nb[3]), sel4(t, D.m[0], D.m[1], D.m[2], D.m[3]),
                    sel4(t, D.s1[0], D.s1[1], D.s1[2], D.s1[3]), sel4(t, D.s2[0], D.s2[1], D.s2[2], D.s2[3]));
}

// Address of slot 0 of key k's bucket in subtable t.
__host__ __device__ inline uint64_t *bucket_ptr(const Desc &D, uint32_t k, uint32_t t)
{
    uint64_t *base = sel4(t, D.sub[0], D.sub[1], D.sub[2], D.sub[3]);
    return base + (uint64_t)bucket_of(D, k, t) * kSlots;
}

// Bucket regions for ordering an insert batch: region r of the table is slice r of every
// subtable, slice = ceil(nb[t] / R) buckets, so region(t, b) = b / slice[t] < R.
struct Regions {
    uint32_t R;
    uint32_t m[kSubtables], s1[kSubtables], s2[kSubtables];  // FastDiv of slice[t]
};

// Region of the bucket a key is first tried in (odd keys start in pos2, cuckoo.cu:336-341).
__host__ __device__ inline uint32_t first_region(const Desc &D, const Regions &M, uint32_t k)
{
    const uint32_t p = pair_of(k);
    const uint32_t t = (k & 1u) ? pos2(p) : pos1(p);
    const uint32_t b = bucket_of(D, k, t);
    return fast_div(b, sel4(t, M.m[0], M.m[1], M.m[2], M.m[3]), sel4(t, M.s1[0], M.s1[1], M.s1[2], M.s1[3]),
                    sel4(t, M.s2[0], M.s2[1], M.s2[2], M.s2[3]));
}

inline Regions make_regions(const Desc &D, uint32_t R)
{
    Regions M{};
    M.R = R;
    for (int t = 0; t < kSubtables; ++t) {
        const uint32_t slice = (D.nb[t] + R - 1) / R;
        FastDiv f = make_fastdiv(slice);
        M.m[t] = f.m;
        M.s1[t] = f.s1;
        M.s2[t] = f.s2;
    }
    return M;
}

inline void desc_set(Desc &D, int t, uint64_t *sub, uint32_t nb)
{
    FastDiv f = make_fastdiv(nb);
    D.sub[t] = sub;
    D.nb[t] = nb;
    D.m[t] = f.m;
    D.s1[t] = f.s1;
    D.s2[t] = f.s2;
}

}  // namespace dyc
