"""File formats on the CPU: raw uint32 dataset files (the reference drivers' input,
GPU/dyhash/Main.cpp:29-59) and the checkpoint header layout."""
import struct

import numpy as np
import pytest

from databases_gpu_amd import io


def test_dataset_round_trip_and_partial_read(tmp_path):
    keys = np.random.default_rng(1).integers(1, 2**32 - 1, size=10_001, dtype=np.uint64).astype(np.uint32)
    p = tmp_path / "keys.dat"
    io.write_keys(p, keys)
    assert p.stat().st_size == keys.size * 4
    assert np.array_equal(io.read_keys(p), keys)
    assert np.array_equal(io.read_keys(p, 100), keys[:100])
    # little-endian on disk whatever the host
    assert p.read_bytes()[:4] == struct.pack("<I", int(keys[0]))
    with pytest.raises(ValueError):
        io.read_keys(p, keys.size + 1)
    bad = tmp_path / "bad.dat"
    bad.write_bytes(b"\x01\x02\x03")
    with pytest.raises(ValueError):
        io.read_keys(bad)


def test_checkpoint_header_layout(tmp_path):
    # byte offsets of the 128-byte header: magic 0, version 8, subtables 12, buckets[6] 16,
    # count 40, pending 48, checksum[6] 56, pending checksum 104, header check 112, reserved 120
    raw = bytearray(128)
    raw[0:8] = b"DYCKPT01"
    struct.pack_into("<II", raw, 8, 2, 5)
    struct.pack_into("<5I", raw, 16, 6, 6, 12, 6, 6)
    struct.pack_into("<QQ", raw, 40, 77, 3)
    struct.pack_into("<5Q", raw, 56, 1, 2, 3, 4, 5)
    struct.pack_into("<QQ", raw, 104, 8, 99)
    fields = dict(version=2, subtables=5, buckets=[6, 6, 12, 6, 6], count=77, pending=3, checksums=[1, 2, 3, 4, 5],
                  pending_checksum=8, header_check=99)
    p = tmp_path / "t.ckpt"
    p.write_bytes(bytes(raw) + b"\0" * 64)
    assert io.checkpoint_header(p) == fields
    p.write_bytes(b"NOTACKPT" + bytes(raw[8:]))
    with pytest.raises(ValueError):
        io.checkpoint_header(p)
    struct.pack_into("<I", raw, 12, 7)  # subtable count out of range
    p.write_bytes(bytes(raw))
    with pytest.raises(ValueError):
        io.checkpoint_header(p)


def test_fnv1a_word_checksum_properties():
    w = np.arange(1, 257, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)
    h = io.fnv1a64_words(w)
    assert io.fnv1a64_words(w[:0]) == 0xCBF29CE484222325  # empty part: the offset basis
    # chaining: hashing in two parts continues from the first part's value
    assert io.fnv1a64_words(w[100:], io.fnv1a64_words(w[:100])) == h
    # any single-word change changes it (multiplication by the odd prime is invertible mod 2^64)
    for i in (0, 17, 255):
        v = w.copy()
        v[i] ^= np.uint64(1)
        assert io.fnv1a64_words(v) != h
