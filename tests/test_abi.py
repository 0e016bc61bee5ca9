"""CPU checks of the drop-in boundary: libdyc.so loads, exports every entry point include/dyc.h
declares, its structs have the C layout the Python binding assumes, and the C++ surfaces compile.
No compute call is made here (there is no GPU in CI)."""
import ctypes
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
CSRC = os.path.join(ROOT, "databases_gpu_amd", "csrc")


def declared_entry_points():
    text = open(os.path.join(INC, "dyc.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dyc_[a-z_0-9]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    import databases_gpu_amd._lib as L

    if not os.path.exists(L.LIB_PATH):
        import databases_gpu_amd._build as B

        B.build()
    return L.load()


def test_library_exports_every_declared_symbol(lib):
    names = declared_entry_points()
    assert len(names) >= 15
    out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(ROOT, "databases_gpu_amd", "libdyc.so")],
                         capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (dyc_\w+)$", out, flags=re.M))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    for n in names:
        assert getattr(lib, n) is not None
    import databases_gpu_amd._lib as L

    assert sorted(L.EXPORTS) == names


def test_abi_version_and_defaults(lib):
    import databases_gpu_amd as D

    assert lib.dyc_abi_version() == 4
    cfg = D.default_config()
    assert cfg.kick_bound == 15 and cfg.rescue_kicks == 128
    assert cfg.lower_bound == 0.5 and cfg.upper_bound == 0.9
    assert cfg.find_lanes == 8 and cfg.block_threads == 256
    assert cfg.min_capacity == int(4 * 1024 * 1024 * 1.2)  # api.cpp:496


def test_errors_without_a_device_are_reported_not_crashed(lib):
    """Argument errors come back as codes with a message (no GPU call is reached)."""
    import databases_gpu_amd._lib as L

    cfg = L.default_config(find_lanes=3)
    h = ctypes.c_void_p()
    rc = lib.dyc_create(ctypes.byref(cfg), 1000, ctypes.byref(h))
    assert rc == 1 and b"lanes" in lib.dyc_last_error()
    assert L.default_config().subtables == 4
    for d in (1, 7):  # subtable counts outside the 2..6 of the reference's pair schemes
        cfg = L.default_config(subtables=d)
        assert lib.dyc_create(ctypes.byref(cfg), 1000, ctypes.byref(h)) == 1 and b"subtables" in lib.dyc_last_error()
        assert lib.dyc_group_create(ctypes.byref(cfg), 1, (ctypes.c_uint64 * 1)(1000), ctypes.byref(h)) == 1
    assert lib.dyc_insert(None, None, None, 0, None) == 1
    assert lib.dyc_insert_counted(None, None, None, 0, None, None) == 1
    assert lib.dyc_find(None, None, None, None, 0, None) == 1
    assert lib.dyc_get_stats(None, None) == 1
    # a subtable holds at most 2^29 - 1 buckets (bucket locations are packed in 29 bits): a hint
    # past it is refused before any device call, for every subtable count and both key widths
    for d in (2, 4, 6):
        cfg = L.default_config(subtables=d)
        too_big = ctypes.c_uint64(d * 16 * (1 << 29))
        assert lib.dyc_create(ctypes.byref(cfg), too_big, ctypes.byref(h)) == 1
        assert b"too large" in lib.dyc_last_error()
        assert lib.dyc_wide_create(ctypes.byref(cfg), too_big, ctypes.byref(h)) == 1
    assert lib.dyc_wide_insert(None, None, None, 0, None) == 1
    assert lib.dyc_wide_find(None, None, None, None, 0, None) == 1
    assert lib.dyc_wide_get_stats(None, None) == 1


def test_struct_layout_matches_c(lib, tmp_path):
    import databases_gpu_amd._lib as L

    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", INC, os.path.join(ROOT, "tests", "cpp", "layout.c"), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    c = [int(x) for x in out[0].split()[1:]]
    s = [int(x) for x in out[1].split()[1:]]
    assert c == [ctypes.sizeof(L.Config), L.Config.lower_bound.offset, L.Config.auto_resize.offset,
                 L.Config.max_blocks.offset, L.Config.min_capacity.offset, L.Config.insert_order.offset,
                 L.Config.region_mib.offset, L.Config.subtables.offset]
    assert s == [ctypes.sizeof(L.Stats), L.Stats.buckets.offset, L.Stats.pending.offset, L.Stats.rejected.offset,
                 L.Stats.lost.offset, L.Stats.subtables.offset]
    r = [int(x) for x in out[2].split()[1:]]
    assert r == [ctypes.sizeof(L.Request), L.Request.keys.offset, L.Request.vals.offset, L.Request.out_flags.offset,
                 L.Request.n.offset]
    assert np.dtype(L.REQUEST_DTYPE).itemsize == r[0]
    sc = [int(x) for x in out[3].split()[1:]]
    assert sc == [ctypes.sizeof(L.ServerConfig), L.ServerConfig.tick_max_requests.offset,
                  L.ServerConfig.item_keys.offset, L.ServerConfig.pool_cap.offset, L.ServerConfig.poll_backoff.offset,
                  L.ServerConfig.resident.offset, L.ServerConfig.resident_idle_ms.offset]
    ss = [int(x) for x in out[4].split()[1:]]
    assert ss == [ctypes.sizeof(L.ServerStats), L.ServerStats.deferred_memory.offset, L.ServerStats.in_flight.offset,
                  L.ServerStats.loop_pauses.offset]
    assert [int(x) for x in out[5].split()[1:]] == [L.Stats.last_max_chain.offset, L.Stats.last_filled.offset]


def _hipcc():
    return shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


def test_server_arena_allocator(tmp_path):
    """The serving loop's HBM arena: random alloc/release against a byte map (no overlap,
    alignment, exact accounting, full coalescing back to one block)."""
    exe = tmp_path / "arena"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", CSRC, os.path.join(ROOT, "tests", "cpp", "arena_check.cpp"),
                    "-o", str(exe)], check=True, capture_output=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "arena ok" in r.stdout


def test_server_argument_errors(lib):
    """Server entry points reject bad arguments with a code and a message before any GPU call."""
    import databases_gpu_amd._lib as L

    h = ctypes.c_void_p()
    hints = (ctypes.c_uint64 * 1)(1000)
    assert lib.dyc_server_create(None, None, 0, hints, ctypes.byref(h)) == 1
    sc = L.ServerConfig()
    assert lib.dyc_server_config_init(ctypes.byref(sc)) == 0
    assert sc.pipeline_depth == 2 and sc.tick_max_keys == 1 << 23 and sc.pool_cap == 65536
    assert sc.one_launch == 1 and sc.spin_cap == 1 << 24
    assert sc.resident == 0 and sc.resident_blocks == 0 and sc.resident_idle_ms == 100
    sc.resident = 2
    assert lib.dyc_server_create(None, ctypes.byref(sc), 1, hints, ctypes.byref(h)) == 1
    assert b"resident" in lib.dyc_last_error()
    sc.resident = 1
    cfg = L.default_config(insert_lanes=8)
    assert lib.dyc_server_create(ctypes.byref(cfg), ctypes.byref(sc), 1, hints, ctypes.byref(h)) == 1
    assert b"insert_lanes" in lib.dyc_last_error()
    sc.resident = 0
    sc.pipeline_depth = 3
    assert lib.dyc_server_create(None, ctypes.byref(sc), 1, hints, ctypes.byref(h)) == 1
    assert b"pipeline_depth" in lib.dyc_last_error()
    assert lib.dyc_server_submit(None, None, None) == 1
    assert lib.dyc_server_drain(None) == 1
    assert lib.dyc_server_get_stats(None, None) == 1


def test_fast_division_is_exact(tmp_path):
    """fast_mod(n, d) == n % d: the bucket index on the device uses it instead of a divide."""
    exe = tmp_path / "fastdiv"
    subprocess.run([_hipcc(), "-O2", "-std=c++17", "-I", CSRC, os.path.join(ROOT, "tests", "cpp", "fastdiv_check.cpp"),
                    "-o", str(exe)], check=True, capture_output=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert " bad 0" in r.stdout


def test_nth_set_bit16_is_exact(tmp_path):
    """nth_set_bit16(m, r): the slot of the bulk fill's r-th placement into a bucket with free-slot
    mask m, checked for all 2^16 masks and every rank against clearing the lowest bit r times."""
    exe = tmp_path / "bits"
    subprocess.run([_hipcc(), "-O2", "-std=c++17", "-I", CSRC, os.path.join(ROOT, "tests", "cpp", "bits_check.cpp"),
                    "-o", str(exe)], check=True, capture_output=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "checked 524288 bad 0" in r.stdout


def test_fill_rule_keeps_subtable_shares_equal_and_favours_pos1(oracle_lib, tmp_path):
    """The bulk fill's first-round rule (Desc::fr, dyc_device.hpp, compiled for the host here):
    per pair, all keys to pos1, all to pos2, or the start rule (odd keys pos2, cuckoo.cu:336-341).
    Over the oracle's pair tables every subtable gets the same share of the keys, and the share
    sent to pos1 equals the best an exhaustive search over all 3^np rules finds (d <= 5) or a
    pruned search's (d = 6); the start rule alone sends half."""
    import itertools

    import oracle as O

    exe = tmp_path / "fill_rule_check"
    subprocess.run([_hipcc(), "-O2", "-std=c++17", "-I", CSRC, os.path.join(ROOT, "tests", "cpp", "fill_rule_check.cpp"),
                    "-o", str(exe)], check=True, capture_output=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, check=True)
    rules = {int(a): int(b) for a, b in (line.split() for line in r.stdout.split("\n") if line.strip())}
    import fillrule

    for d in range(2, 7):
        assert rules[d] == fillrule.rule(d), d  # the tests' restatement picks the same rule
        pairs = O.pair_table(d)
        halves = {1: 2, 0: 1, 2: 0}  # code -> halves of the pair's keys sent to pos1
        load = [0] * d
        to_pos1 = 0
        for p, (a, b) in enumerate(pairs):
            x = halves[(rules[d] >> (2 * p)) & 3]
            load[a] += x
            load[b] += 2 - x
            to_pos1 += x
        assert len(set(load)) == 1, (d, load)  # equal shares
        if d <= 5:  # exhaustive
            best = 0
            for xs in itertools.product((2, 1, 0), repeat=len(pairs)):
                ld = [0] * d
                for (a, b), x in zip(pairs, xs):
                    ld[a] += x
                    ld[b] += 2 - x
                if len(set(ld)) == 1:
                    best = max(best, sum(xs))
            assert to_pos1 == best, (d, to_pos1, best)
        else:
            assert to_pos1 * 10 == 7 * 2 * len(pairs)  # 0.7 of the keys
        assert to_pos1 > len(pairs)  # more than the start rule's half


def test_device_header_hashes_match_the_oracle(oracle_lib, tmp_path):
    """The kernels' per-key arithmetic (dyc_device.hpp, compiled for the host here) equals the
    oracle's on edge and random keys, for every subtable count. The device enumerates each
    pair table; the oracle states the reference's per-d formulas, so this also checks one
    representation against the other."""
    import numpy as np

    import oracle as O

    src = tmp_path / "dump.cpp"
    src.write_text(r'''
#include <cstdio>
#include "dyc_device.hpp"
int main() {
    unsigned k;
    const unsigned nb[6] = {2u, 36424u, 72818u, 4660218u, 1000003u, 98u};
    while (scanf("%u", &k) == 1) {
        for (unsigned d = 2; d <= 6; ++d) {
            dyc::Desc D{};
            if (!dyc::desc_init(D, d)) return 1;
            for (unsigned t = 0; t < d; ++t) dyc::desc_set(D, (int)t, nullptr, nb[t]);
            const unsigned p = dyc::pair_of(D, k);
            printf("%u %u %u %u %u %u %u %u %u %u", dyc::h1(k), dyc::h2(k), dyc::h3(k), dyc::h4(k), dyc::h5(k), p,
                   dyc::pos1(D, p), dyc::pos2(D, p), dyc::other_pos(D, p, 100u), dyc::rescue_lane(k, 7));
            for (unsigned t = 0; t < d; ++t) printf(" %u", dyc::bucket_of(D, k, t));
            printf("\n");
        }
    }
}
''')
    exe = tmp_path / "dump"
    subprocess.run([_hipcc(), "-O2", "-std=c++17", "-I", CSRC, str(src), "-o", str(exe)], check=True,
                   capture_output=True)
    rng = np.random.default_rng(3)
    keys = np.concatenate([np.array([1, 2, 11, 0x7FFFFFFF, 0x80000000, 0xFFFFFFFE], dtype=np.uint32),
                           rng.integers(1, 2**32 - 1, size=1500, dtype=np.uint64).astype(np.uint32)])
    out = subprocess.run([str(exe)], input="\n".join(map(str, keys.tolist())), capture_output=True, text=True,
                         check=True).stdout
    lines = out.strip().split("\n")
    assert len(lines) == 5 * keys.size
    L = O.lib()
    nbs = [2, 36424, 72818, 4660218, 1000003, 98]
    for i, k in enumerate(keys.tolist()):
        for j, d in enumerate(range(2, 7)):
            got = [int(x) for x in lines[5 * i + j].split()]
            p = L.orc_pair_d(k, d)
            want = [L.orc_h1(k), L.orc_h2(k), L.orc_h3(k), L.orc_h4(k), L.orc_h5(k), p, L.orc_pos1_d(p, d),
                    L.orc_pos2_d(p, d), L.orc_get_pos_d(p, 100, d), L.orc_rescue_lane(k, 7),
                    *[L.orc_bucket(k, t, nbs[t]) for t in range(d)]]
            assert got == want, (k, d)


def test_counter_block_folds_every_stripe_into_the_base(tmp_path):
    """Per-wave tallies land in one of 64 striped copies of the counter block; the host adds every
    stripe into the base copy once (fold_tallies), except max_chain (word 12), where it keeps the
    largest. A block whose stripe s holds 1000 * (s + 1) + word, over base words 7 * word, must fold
    to the closed-form sums (word 12: the last stripe's 1000 * 64 + 12) and leave the stripes."""
    src = tmp_path / "fold.cpp"
    src.write_text(r'''
#include <cstdio>
#include <vector>
#include "dyc_kernels.hpp"
int main() {
    using namespace dyc;
    static_assert(sizeof(Counters) == 16 * sizeof(unsigned long long), "16 counter words");
    std::vector<unsigned long long> w(kCounterWords, 0);
    for (int i = 0; i < 16; ++i) w[i] = 7ull * i;
    for (int s = 0; s < kTallyStripes; ++s)
        for (int i = 0; i < 16; ++i) w[kTallyStrideWords * (s + 1) + i] = 1000ull * (s + 1) + i;
    fold_tallies(w.data());
    const Counters &c = *reinterpret_cast<const Counters *>(w.data());
    int bad = 0;
    const unsigned long long S = kTallyStripes;
    for (int i = 0; i < 16; ++i)
        bad += w[i] != (i == kMaxChainWord ? 1000ull * S + i : 7ull * i + 1000ull * S * (S + 1) / 2 + S * i);
    for (int s = 0; s < kTallyStripes; ++s)
        for (int i = 0; i < 16; ++i) bad += w[kTallyStrideWords * (s + 1) + i] != 1000ull * (s + 1) + i;
    bad += c.deleted != w[4] || c.placed != w[11] || c.max_chain != w[12] || kMaxChainWord != 12;
    printf("words %d stripes %d stride %d bad %d\n", kCounterWords, kTallyStripes, kTallyStrideWords, bad);
    return bad != 0;
}
''')
    exe = tmp_path / "fold"
    subprocess.run([_hipcc(), "-O2", "-std=c++17", "-I", CSRC, str(src), "-o", str(exe)], check=True,
                   capture_output=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.split() == ["words", "2080", "stripes", "64", "stride", "32", "bad", "0"]


def test_cpp_surfaces_compile(tmp_path):
    """include/dyc_hashapi.hpp and the reference-test driver build against the C ABI."""
    obj = tmp_path / "driver.o"
    subprocess.run([_hipcc(), "-O1", "-std=c++17", "-I", INC, "-c", os.path.join(ROOT, "tests", "cpp", "hashapi_driver.cpp"),
                    "-o", str(obj)], check=True, capture_output=True)
    assert obj.exists()
