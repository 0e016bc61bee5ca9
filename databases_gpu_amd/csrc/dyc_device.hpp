// dyc_device.hpp -- DyCuckoo table descriptor and per-key arithmetic shared by every kernel.
//
// Layout (the reference's, cuckoo.h:13-16,30-34): d subtables (d = 4 by default, 2..6 allowed),
// each an array of 128-byte buckets of 16 uint64 slots; slot = (uint64)key << 32 | value,
// empty = 0. Keys 0 (empty marker) and 0xFFFFFFFF are reserved.
//
// All functions are __host__ __device__ so the host code computes the same numbers the kernels
// do (dump checks, the fast-division parameters).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <array>

namespace dyc {

constexpr int kDefaultSubtables = 4;           // dy_hash.h:32 TABLE_NUM
constexpr int kMinSubtables = 2;               // the n-choose-2 variants, diff_hash_num_pos_fun.cpp:14-188
constexpr int kMaxSubtables = 6;
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

// Subtable t's hash reduced mod PRIME_uint: h_{t % 4} (cuckoo.cu:76-90, `switch(num_table%4)`,
// so subtables 4 and 5 reuse h1 and h2). The four hashes are evaluated branch-free and
// selected: t differs lane to lane, and a divergent switch would run all four paths under
// masks anyway.
__host__ __device__ inline uint32_t hash_t(uint32_t k, uint32_t t)
{
    const uint32_t a = h1(k), b = h2(k), c = h3(k), d = h4(k);
    const uint32_t u = t & 3u;
    const uint32_t x = u == 0 ? a : (u == 1 ? b : (u == 2 ? c : d));
    return x % kPrime;
}

// Victim lane of the rescue phase that follows the reference kick bound: the top 4 bits of a
// multiplicative hash of the in-hand key and the iteration. (The reference drops the entry at
// its kick bound, cuckoo.cu:422-431.)
__host__ __device__ inline uint32_t rescue_lane(uint32_t k, uint32_t it)
{
    return (k * 0x9E3779B1u + it * 0x632BE5ABu) >> 28;
}

// Position of the set bit of rank r (0 = lowest) in a 16-bit slot mask m, for r < popcount(m):
// four halving steps, each keeping the half that holds it (no loop, no divergence).
__host__ __device__ inline uint32_t nth_set_bit16(uint32_t m, uint32_t r)
{
    uint32_t pos = 0, c = (uint32_t)__builtin_popcount(m & 0xffu);
    if (r >= c) r -= c, m >>= 8, pos += 8;
    c = (uint32_t)__builtin_popcount(m & 0xfu);
    if (r >= c) r -= c, m >>= 4, pos += 4;
    c = (uint32_t)__builtin_popcount(m & 0x3u);
    if (r >= c) r -= c, m >>= 2, pos += 2;
    return pos + (r >= (m & 1u) ? 1u : 0u);
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
// live in SGPRs; per-lane selection by subtable is a v_cndmask chain over the kMaxSubtables
// entries (entries at and past d are never selected).
struct Desc {
    uint64_t *sub[kMaxSubtables];  // bucket arrays, nb[t] * 16 slots each
    uint32_t nb[kMaxSubtables];    // buckets per subtable
    uint32_t m[kMaxSubtables];     // FastDiv multipliers for nb[t]
    uint32_t sh[kMaxSubtables];    // FastDiv shifts for nb[t]: s1 | s2 << 8
    uint32_t d;                    // subtables in use, kMinSubtables..kMaxSubtables
    uint32_t np;                   // pairs: d (d - 1) / 2, and 2 for d = 2
    uint32_t pm, psh;              // FastDiv of np (multiplier, s1 | s2 << 8)
    uint64_t p1, p2;               // pos1 / pos2 of pair p in bits [4p, 4p + 4)
    uint32_t fr;                   // bulk fill, first round: bits [2p, 2p + 2) say where pair p's
                                   // keys go (fill_target); 0 (the start rule) if never set
};

// Values are read at constant indices first and then selected: a conditional over the array
// elements themselves would select an address, and a descriptor indexed that way is kept in
// scratch memory.
template <typename T>
__host__ __device__ inline T selt(uint32_t t, const T (&a)[kMaxSubtables])
{
    static_assert(kMaxSubtables == 6, "selection chain below lists six entries");
    const T a0 = a[0], a1 = a[1], a2 = a[2], a3 = a[3], a4 = a[4], a5 = a[5];
    T x = t == 1 ? a1 : a0;
    x = t == 2 ? a2 : x;
    x = t == 3 ? a3 : x;
    x = t == 4 ? a4 : x;
    return t == 5 ? a5 : x;
}

__host__ __device__ inline uint32_t fast_mod_sh(uint32_t n, uint32_t d, uint32_t m, uint32_t sh)
{
    return fast_mod(n, d, m, sh & 0xffu, sh >> 8);
}

// ------------------------------------------------------------------------- pair selection
// Two-layer scheme (cuckoo.cu:190-213,310; the other subtable counts in
// diff_hash_num_pos_fun.cpp:14-188): a key may live only in the two subtables pos1/pos2 of its
// pair. The pair is h5(k) % C(d,2), except for d = 2 where it is k & 1
// (diff_hash_num_pos_fun.cpp:16-19). For d = 4 the pairs 0..5 are (1,2) (0,1) (2,3) (0,2)
// (1,3) (0,3).
__host__ __device__ inline uint32_t pair_of(const Desc &D, uint32_t k)
{
    const uint32_t src = D.d == 2u ? k : h5(k);
    return fast_mod_sh(src, D.np, D.pm, D.psh);
}
__host__ __device__ inline uint32_t pos1(const Desc &D, uint32_t p) { return (uint32_t)(D.p1 >> (4u * p)) & 15u; }
__host__ __device__ inline uint32_t pos2(const Desc &D, uint32_t p) { return (uint32_t)(D.p2 >> (4u * p)) & 15u; }
__host__ __device__ inline uint32_t other_pos(const Desc &D, uint32_t p, uint32_t pre)
{
    const uint32_t a = pos1(D, p);
    return pre == a ? pos2(D, p) : a;
}

// The reference victim vote (cuckoo.cu:396-401) takes the first slot whose resident's pair
// index differs from the table number being written (a pair compared with a table, kept for
// layout parity); for d = 2 the vote is over every slot (diff_hash_num_pos_fun.cpp:3-6), so
// a full bucket always gives up slot 0.
__host__ __device__ inline bool vote_victim(const Desc &D, uint32_t resident_key, uint32_t t)
{
    return D.d == 2u || pair_of(D, resident_key) != t;
}

__host__ __device__ inline uint32_t bucket_of(const Desc &D, uint32_t k, uint32_t t)
{
    return fast_mod_sh(hash_t(k, t), selt(t, D.nb), selt(t, D.m), selt(t, D.sh));
}

// One subtable as code that works inside it alone sees it: its arrays, bucket count and
// FastDiv constants, and its index (which selects the hash).
struct SubRef {
    uint64_t *keys;  // bucket array, 16 slots per bucket (key | value entries, or wide keys)
    uint64_t *vals;  // value array parallel to it (wide tables), else null
    uint32_t nb, m, sh, t;
};

__host__ __device__ inline SubRef sub_ref(const Desc &D, uint32_t t, uint64_t *vals = nullptr)
{
    return SubRef{selt(t, D.sub), vals, selt(t, D.nb), selt(t, D.m), selt(t, D.sh), t};
}

// bucket_of for a subtable that every lane of the wave shares: a scalar branch evaluates the
// one hash that it uses instead of all four.
__device__ inline uint32_t bucket_of_uniform(const SubRef &S, uint32_t k)
{
    uint32_t x;
    switch (S.t & 3u) {
    case 0: x = h1(k) % kPrime; break;
    case 1: x = h2(k) % kPrime; break;
    case 2: x = h3(k); break;
    default: x = h4(k) % kPrime; break;
    }
    return fast_mod_sh(x, S.nb, S.m, S.sh);
}

__device__ inline uint32_t bucket_of_uniform(const Desc &D, uint32_t k, uint32_t t)
{
    return bucket_of_uniform(sub_ref(D, t), k);
}

// Address of slot 0 of key k's bucket in subtable t.
__host__ __device__ inline uint64_t *bucket_ptr(const Desc &D, uint32_t k, uint32_t t)
{
    return selt(t, D.sub) + (uint64_t)bucket_of(D, k, t) * kSlots;
}

// Bucket regions for ordering an insert batch: region r of the table is slice r of every
// subtable, slice = ceil(nb[t] / R) buckets, so region(t, b) = b / slice[t] < R.
struct Regions {
    uint32_t R;
    uint32_t m[kMaxSubtables], sh[kMaxSubtables];  // FastDiv of slice[t]
};

// Region of the bucket a key is first tried in (odd keys start in pos2, cuckoo.cu:336-341).
__host__ __device__ inline uint32_t first_region(const Desc &D, const Regions &M, uint32_t k)
{
    const uint32_t p = pair_of(D, k);
    const uint32_t t = (k & 1u) ? pos2(D, p) : pos1(D, p);
    const uint32_t b = bucket_of(D, k, t);
    const uint32_t sh = selt(t, M.sh);
    return fast_div(b, selt(t, M.m), sh & 0xffu, sh >> 8);
}

inline uint32_t pack_shift(const FastDiv &f) { return f.s1 | (f.s2 << 8); }

inline Regions make_regions(const Desc &D, uint32_t R)
{
    Regions M{};
    M.R = R;
    for (uint32_t t = 0; t < D.d; ++t) {
        const uint32_t slice = (D.nb[t] + R - 1) / R;
        FastDiv f = make_fastdiv(slice);
        M.m[t] = f.m;
        M.sh[t] = pack_shift(f);
    }
    return M;
}

// The (pos1, pos2) table of every supported subtable count, pair by pair.
//   d = 2: (0,1) (1,0)                     diff_hash_num_pos_fun.cpp:14-29 (pos1 = p, pos2 = p ^ 1)
//   d = 3: (0,1) (1,2) (0,2)               the shift form, diff_hash_num_pos_fun.cpp:37-55
//   d = 4: (1,2) (0,1) (2,3) (0,2) (1,3) (0,3)                         cuckoo.cu:189-213
//   d = 5: the d = 4 pairs, then (2,4) (0,4) (1,4) (3,4)       diff_hash_num_pos_fun.cpp:119-146
//   d = 6: every pair (a,b), a < b, in lexicographic order     diff_hash_num_pos_fun.cpp:152-188
inline bool pair_scheme(uint32_t d, uint32_t &np, uint64_t &p1, uint64_t &p2)
{
    static const uint8_t k2[][2] = {{0, 1}, {1, 0}};
    static const uint8_t k3[][2] = {{0, 1}, {1, 2}, {0, 2}};
    static const uint8_t k4[][2] = {{1, 2}, {0, 1}, {2, 3}, {0, 2}, {1, 3}, {0, 3}};
    static const uint8_t k5[][2] = {{1, 2}, {0, 1}, {2, 3}, {0, 2}, {1, 3}, {0, 3}, {2, 4}, {0, 4}, {1, 4}, {3, 4}};
    uint8_t k6[15][2];
    uint32_t q = 0;
    for (uint8_t a = 0; a < 6; ++a)
        for (uint8_t b = a + 1; b < 6; ++b) {
            k6[q][0] = a;
            k6[q][1] = b;
            ++q;
        }
    const uint8_t(*tab)[2] = nullptr;
    switch (d) {
    case 2: tab = k2; np = 2; break;
    case 3: tab = k3; np = 3; break;
    case 4: tab = k4; np = 6; break;
    case 5: tab = k5; np = 10; break;
    case 6: tab = k6; np = 15; break;
    default: return false;
    }
    p1 = p2 = 0;
    for (uint32_t i = 0; i < np; ++i) {
        p1 |= (uint64_t)tab[i][0] << (4 * i);
        p2 |= (uint64_t)tab[i][1] << (4 * i);
    }
    return true;
}

// The bucket the bulk fill's first round sends key k of pair p to. The eviction insert starts a
// key by the reference's start rule (odd keys in pos2, even ones in pos1, cuckoo.cu:336-341),
// which gives every subtable the same share of the keys but puts only half of them in pos1, the
// bucket a find reads first. Where the fill puts a key is its own choice: per pair, Desc::fr
// picks 1 = always pos1, 2 = always pos2, 0 = the start rule, so that the subtables' shares stay
// equal and the most keys go to pos1 (fill_rule).
__host__ __device__ inline uint32_t fill_target(const Desc &D, uint32_t k, uint32_t p)
{
    const uint32_t r = (D.fr >> (2u * p)) & 3u;
    return (r == 2u || (r == 0u && (k & 1u))) ? pos2(D, p) : pos1(D, p);
}

// A fill round's subtable for keys of pair p = idx & 15 and parity idx >> 4 (the index of a
// 32-entry table the pack kernels keep in LDS, so a key's subtable costs one LDS read): round 1
// fill_target, round 2 the other bucket of the pair.
__host__ __device__ inline uint32_t fill_round_sub(const Desc &D, uint32_t idx, bool round2)
{
    const uint32_t p = idx & 15u;
    if (p >= D.np) return 0;
    const uint32_t t = fill_target(D, idx >> 4, p);
    return round2 ? other_pos(D, p, t) : t;
}

// Depth-first search for Desc::fr over the np pairs (3^np choices, pruned): each pair's keys go
// all to pos1, half and half, or all to pos2, no subtable may exceed an equal share (2 np / d
// halves of a pair's keys; at the end every share is then equal), and the most halves go to
// pos1. The shares in pos1 for d = 2..6: 1, 2/3, 2/3, 0.7, 0.7 (the start rule: 1/2).
struct FillRuleSearch {
    uint32_t np = 0, share = 0;
    uint32_t a[16] = {}, b[16] = {};  // pos1, pos2 of each pair
    uint32_t load[kMaxSubtables] = {};
    uint32_t cur = 0, best = 0, best_score = 0;
    bool found = false;
    void go(uint32_t i, uint32_t score)
    {
        if (found && score + 2 * (np - i) <= best_score) return;
        if (i == np) {
            found = true;
            best = cur;
            best_score = score;
            return;
        }
        static const uint32_t kHalves[3] = {2, 1, 0}, kCode[3] = {1, 0, 2};  // pos1 first
        for (int c = 0; c < 3; ++c) {
            const uint32_t x = kHalves[c];
            load[a[i]] += x;
            load[b[i]] += 2 - x;
            if (load[a[i]] <= share && load[b[i]] <= share) {
                cur = (cur & ~(3u << (2 * i))) | (kCode[c] << (2 * i));
                go(i + 1, score + x);
            }
            load[a[i]] -= x;
            load[b[i]] -= 2 - x;
        }
    }
};

inline uint32_t fill_rule_search(uint32_t d)
{
    FillRuleSearch S;
    uint64_t p1 = 0, p2 = 0;
    if (!pair_scheme(d, S.np, p1, p2) || S.np > 15 || (2 * S.np) % d) return 0;
    S.share = 2 * S.np / d;
    for (uint32_t p = 0; p < S.np; ++p) {
        S.a[p] = (uint32_t)(p1 >> (4 * p)) & 15u;
        S.b[p] = (uint32_t)(p2 >> (4 * p)) & 15u;
    }
    S.go(0, 0);
    return S.found ? S.best : 0u;
}

// Desc::fr for d subtables (searched once per d).
inline uint32_t fill_rule(uint32_t d)
{
    static const std::array<uint32_t, kMaxSubtables + 1> rules = [] {
        std::array<uint32_t, kMaxSubtables + 1> r{};
        for (uint32_t x = (uint32_t)kMinSubtables; x <= (uint32_t)kMaxSubtables; ++x) r[x] = fill_rule_search(x);
        return r;
    }();
    return d <= (uint32_t)kMaxSubtables ? rules[d] : 0u;
}

// An empty descriptor for d subtables (false if d is unsupported).
inline bool desc_init(Desc &D, uint32_t d)
{
    D = Desc{};
    uint32_t np;
    if (!pair_scheme(d, np, D.p1, D.p2)) return false;
    const FastDiv f = make_fastdiv(np);
    D.d = d;
    D.np = np;
    D.pm = f.m;
    D.psh = pack_shift(f);
    D.fr = fill_rule(d);
    return true;
}

inline void desc_set(Desc &D, int t, uint64_t *sub, uint32_t nb)
{
    FastDiv f = make_fastdiv(nb);
    D.sub[t] = sub;
    D.nb[t] = nb;
    D.m[t] = f.m;
    D.sh[t] = pack_shift(f);
}

}  // namespace dyc
