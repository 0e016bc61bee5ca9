"""MI355X-native DyCuckoo: a dynamic multi-subtable cuckoo hash table on gfx950.

The hot path (insert / find / erase / single-subtable resize) runs in hand-written HIP kernels
compiled into ``libdyc.so``; its C ABI is ``include/dyc.h``. This package is the Python host
side over that ABI. There is no CPU fallback: if ``libdyc.so`` is missing, using the table
raises.
"""
from ._lib import SLOTS, SUBTABLES, Config, DycError, Stats, default_config, load  # noqa: F401
from .server import DyCuckooServer  # noqa: F401
from .table import DyCuckoo, DyCuckooGroup, DyCuckooWide, DyCuckooWideGroup, HashAPI  # noqa: F401

__all__ = ["DyCuckoo", "DyCuckooGroup", "DyCuckooServer", "DyCuckooWide", "DyCuckooWideGroup", "HashAPI", "DycError", "Config", "Stats", "default_config", "load", "SLOTS",
           "SUBTABLES"]
