"""CPU sanitizer builds (ASan + UBSan, TSan) of the host-side code that runs threads or manages
memory on the host: the oracle (including the multithreaded CPU baseline, which shares a table
image through 64-bit CAS), the library's worker threads (the deferred-free helper and the copy
pool of the host-array calls, dyc_threads.hpp) and the serving arena allocator. Each program
checks its own results; the sanitizers check memory safety, undefined behaviour and data races.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "databases_gpu_amd", "csrc")
ORACLE = os.path.join(ROOT, "oracle")
CPP = os.path.join(ROOT, "tests", "cpp")

ASAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer"]
TSAN = ["-fsanitize=thread"]
REPORTS = ("ERROR: AddressSanitizer", "ERROR: LeakSanitizer", "runtime error:", "WARNING: ThreadSanitizer")

PROGRAMS = {
    "oracle": (["gcc", "-std=c11", "-pthread", "-I", ORACLE, os.path.join(CPP, "oracle_stress.c"),
                os.path.join(ORACLE, "dyc_oracle.c"), "-lm"], "oracle stress ok"),
    "threads": (["g++", "-std=c++17", "-pthread", "-I", CSRC, os.path.join(CPP, "threads_check.cpp")], "threads ok"),
    "arena": (["g++", "-std=c++17", "-I", CSRC, os.path.join(CPP, "arena_check.cpp")], "arena ok"),
}


@pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("g++") is None, reason="no gcc/g++")
@pytest.mark.parametrize("name,san", [("oracle", "asan"), ("oracle", "tsan"), ("threads", "asan"),
                                      ("threads", "tsan"), ("arena", "asan")])
def test_sanitizer_build_is_clean(tmp_path, name, san):
    cmd, ok_line = PROGRAMS[name]
    exe = tmp_path / f"{name}_{san}"
    flags = ASAN if san == "asan" else TSAN
    subprocess.run([cmd[0], "-O1", "-g", *flags, *cmd[1:], "-o", str(exe)], check=True, capture_output=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", TSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=300)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert ok_line in r.stdout
    assert not any(s in out for s in REPORTS), out[-4000:]
