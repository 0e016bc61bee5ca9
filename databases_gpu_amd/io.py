"""File formats around the table.

* Dataset files as the reference's drivers read them (GPU/dyhash/Main.cpp:29-59): a raw array of
  little-endian uint32 keys with no header; the drivers store value i + 1 for key i
  (Main.cpp:112-116). ``read_keys`` memory-maps one, ``ingest`` inserts it through the pipelined
  host-pointer path without loading it whole.
* Table checkpoints (``DyCuckoo.checkpoint`` / ``DyCuckoo.restore``, dyc_checkpoint_write/read
  in include/dyc.h); ``checkpoint_header`` parses a checkpoint's 128-byte header.
"""
from __future__ import annotations

import os
import struct

import numpy as np

CKPT_MAGIC = b"DYCKPT01"
CKPT_VERSION = 2
# magic, version, subtables d, buckets[6], count, pending, checksum[6] (subtable images),
# pending-list checksum, header_check, reserved[8]; entries at and past d are 0
_CKPT_HEADER = struct.Struct("<8sII6IQQ6QQQ8s")
assert _CKPT_HEADER.size == 128


def read_keys(path: str, count: int | None = None) -> np.ndarray:
    """The uint32 keys of a raw dataset file, memory-mapped (first `count` keys if given)."""
    size = os.path.getsize(path)
    if size % 4:
        raise ValueError(f"{path}: {size} bytes is not a whole number of uint32 keys")
    n = size // 4 if count is None else int(count)
    if n > size // 4:
        raise ValueError(f"{path} holds {size // 4} keys, {n} requested")
    return np.memmap(path, dtype="<u4", mode="r", shape=(n,)) if n else np.zeros(0, dtype=np.uint32)


def write_keys(path: str, keys) -> None:
    """Write keys as a raw little-endian uint32 dataset file."""
    np.ascontiguousarray(np.asarray(keys, dtype="<u4")).tofile(path)


def ingest(table, path: str, count: int | None = None, chunk: int = 1 << 24) -> int:
    """Insert the keys of a dataset file with the reference drivers' values (i + 1 for key i),
    `chunk` keys per host-pointer call. Returns the number of keys read."""
    keys = read_keys(path, count)
    n = keys.size
    for lo in range(0, n, chunk):
        hi = min(n, lo + chunk)
        vals = np.arange(lo + 1, hi + 1, dtype=np.uint64).astype(np.uint32)
        table.insert_host(np.ascontiguousarray(keys[lo:hi]), vals)
    return n


def checkpoint_header(path: str) -> dict:
    """The fields of a checkpoint file's header (no validation beyond the magic)."""
    with open(path, "rb") as f:
        raw = f.read(_CKPT_HEADER.size)
    if len(raw) != _CKPT_HEADER.size:
        raise ValueError(f"{path}: truncated checkpoint header")
    f = _CKPT_HEADER.unpack(raw)
    magic, version, d = f[0], f[1], f[2]
    if magic != CKPT_MAGIC:
        raise ValueError(f"{path}: not a DyCuckoo checkpoint")
    if not 2 <= d <= 6:
        raise ValueError(f"{path}: subtable count {d} out of range")
    buckets, (count, pending), sums = f[3:9], f[9:11], f[11:17]
    return {"version": version, "subtables": d, "buckets": list(buckets[:d]), "count": count,
            "pending": pending, "checksums": list(sums[:d]), "pending_checksum": f[17], "header_check": f[18]}


def fnv1a64_words(words: np.ndarray, h: int = 0xCBF29CE484222325) -> int:
    """FNV-1a over 8-byte words, the checksum of each part of a checkpoint (pure Python: for
    tests and small images)."""
    prime = 0x100000001B3
    for w in np.asarray(words, dtype=np.uint64).tolist():
        h = ((h ^ w) * prime) & 0xFFFFFFFFFFFFFFFF
    return h
