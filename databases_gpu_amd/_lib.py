"""ctypes binding of libdyc.so (the C ABI declared in include/dyc.h).

torch is imported before the library is loaded so that libdyc's dependency on the HIP runtime
(soname libamdhip64.so.7) resolves to the runtime torch already loaded: one process, one HIP
runtime.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdyc.so")

SUBTABLES = 4          # default subtable count (TABLE_NUM)
MAX_SUBTABLES = 6
SUBTABLE_COUNTS = (2, 3, 4, 5, 6)
SLOTS = 16

DYC_OK = 0
ERRORS = {
    1: "DYC_ERR_INVALID",
    2: "DYC_ERR_HIP",
    3: "DYC_ERR_OOM",
    4: "DYC_ERR_PENDING",
    5: "DYC_ERR_ODD_BUCKETS",
    6: "DYC_ERR_LOST",
    7: "DYC_ERR_CORRUPT",
}
DYC_ERR_PENDING = 4
DYC_ERR_ODD_BUCKETS = 5

# every entry point include/dyc.h declares
EXPORTS = (
    "dyc_abi_version", "dyc_last_error", "dyc_config_init", "dyc_create", "dyc_destroy",
    "dyc_insert", "dyc_find", "dyc_erase", "dyc_resize_up", "dyc_resize_down", "dyc_get_stats",
    "dyc_count_occupied", "dyc_dump", "dyc_load", "dyc_pending", "dyc_clear", "dyc_probe_count",
    "dyc_get_config", "dyc_set_config", "dyc_upsert", "dyc_insert_host", "dyc_find_host", "dyc_erase_host",
    "dyc_checkpoint_write", "dyc_checkpoint_read", "dyc_group_create", "dyc_group_destroy", "dyc_group_get_launch", "dyc_group_set_launch",
    "dyc_group_insert", "dyc_group_find", "dyc_group_erase", "dyc_group_insert_per_table",
    "dyc_group_find_per_table", "dyc_group_table_stats", "dyc_group_dump", "dyc_group_pending",
    "dyc_group_clear",
    "dyc_server_config_init", "dyc_server_create", "dyc_server_destroy", "dyc_server_submit", "dyc_server_poll",
    "dyc_server_drain", "dyc_server_replay", "dyc_server_get_stats", "dyc_server_table_stats", "dyc_server_dump",
    "dyc_wide_create", "dyc_wide_destroy", "dyc_wide_insert", "dyc_wide_upsert", "dyc_wide_find", "dyc_wide_erase",
    "dyc_wide_resize_up", "dyc_wide_resize_down", "dyc_wide_get_stats", "dyc_wide_clear", "dyc_wide_dump",
    "dyc_wide_load", "dyc_wide_pending", "dyc_insert_noalloc", "dyc_insert_async", "dyc_sync",
    "dyc_insert_counted", "dyc_group_plan_create", "dyc_group_plan_run", "dyc_group_plan_destroy",
    "dyc_wide_group_create", "dyc_wide_group_destroy", "dyc_wide_group_insert", "dyc_wide_group_find",
    "dyc_wide_group_erase", "dyc_wide_group_table_stats", "dyc_wide_group_dump", "dyc_wide_group_pending",
    "dyc_wide_group_clear", "dyc_wide_group_insert_per_table", "dyc_wide_group_find_per_table",
)

OP_INSERT, OP_FIND, OP_ERASE = 0, 1, 2

NOALLOC_MAX_KEYS = 65536          # DYC_NOALLOC_MAX_KEYS
NOALLOC_SCRATCH_BYTES = 16640     # DYC_NOALLOC_SCRATCH_BYTES


class DycError(RuntimeError):
    def __init__(self, code: int, where: str, msg: str):
        super().__init__(f"{where}: {ERRORS.get(code, code)}: {msg}")
        self.code = code


class Config(ctypes.Structure):
    _fields_ = [
        ("kick_bound", ctypes.c_uint32),
        ("rescue_kicks", ctypes.c_uint32),
        ("max_retries", ctypes.c_uint32),
        ("max_fail_rounds", ctypes.c_uint32),
        ("lower_bound", ctypes.c_double),
        ("upper_bound", ctypes.c_double),
        ("auto_resize", ctypes.c_uint32),
        ("find_lanes", ctypes.c_uint32),
        ("insert_lanes", ctypes.c_uint32),
        ("block_threads", ctypes.c_uint32),
        ("max_blocks", ctypes.c_uint32),
        ("insert_blocks", ctypes.c_uint32),
        ("min_capacity", ctypes.c_uint64),
        ("insert_order", ctypes.c_uint32),
        ("region_mib", ctypes.c_uint32),
        ("subtables", ctypes.c_uint32),
        ("reserved0", ctypes.c_uint32),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("count", ctypes.c_uint64),
        ("capacity", ctypes.c_uint64),
        ("buckets", ctypes.c_uint32 * MAX_SUBTABLES),
        ("pending", ctypes.c_uint64),
        ("kicks", ctypes.c_uint64),
        ("last_kicks", ctypes.c_uint64),
        ("kick_failures", ctypes.c_uint64),
        ("cas_retries", ctypes.c_uint64),
        ("rejected", ctypes.c_uint64),
        ("resizes_up", ctypes.c_uint64),
        ("resizes_down", ctypes.c_uint64),
        ("lost", ctypes.c_uint64),
        ("subtables", ctypes.c_uint32),
        ("last_max_chain", ctypes.c_uint32),
        ("last_filled", ctypes.c_uint64),
    ]

    def as_dict(self) -> dict:
        d = {name: getattr(self, name) for name, _ in self._fields_}
        d["buckets"] = list(self.buckets)[: self.subtables]
        return d


class GroupLaunch(ctypes.Structure):
    _fields_ = [
        ("find_lanes", ctypes.c_uint32),
        ("insert_lanes", ctypes.c_uint32),
        ("find_blocks", ctypes.c_uint32),
        ("insert_blocks", ctypes.c_uint32),
        ("item_keys", ctypes.c_uint32),
        ("kick_bound", ctypes.c_uint32),
        ("insert_waves_per_eu", ctypes.c_uint32),
        ("find_waves_per_eu", ctypes.c_uint32),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class Request(ctypes.Structure):
    _fields_ = [
        ("table", ctypes.c_uint32),
        ("op", ctypes.c_uint32),
        ("keys", ctypes.c_void_p),
        ("vals", ctypes.c_void_p),
        ("out_vals", ctypes.c_void_p),
        ("out_flags", ctypes.c_void_p),
        ("n", ctypes.c_uint64),
    ]


class ServerConfig(ctypes.Structure):
    _fields_ = [
        ("hbm_budget", ctypes.c_uint64),
        ("tick_max_keys", ctypes.c_uint64),
        ("tick_max_requests", ctypes.c_uint32),
        ("pipeline_depth", ctypes.c_uint32),
        ("item_keys", ctypes.c_uint32),
        ("pool_cap", ctypes.c_uint32),
        ("one_launch", ctypes.c_uint32),
        ("spin_cap", ctypes.c_uint32),
        ("poll_backoff", ctypes.c_uint32),
        ("resident", ctypes.c_uint32),
        ("resident_blocks", ctypes.c_uint32),
        ("resident_idle_ms", ctypes.c_uint32),
        ("pad", ctypes.c_uint32),
    ]


class ServerStats(ctypes.Structure):
    _fields_ = [(name, ctypes.c_uint64) for name in (
        "submitted", "completed", "failed", "ticks", "rounds", "keys", "deferred_order", "deferred_memory", "deferred_slots",
        "resizes_up", "resizes_down", "pool_entries", "pool_rounds", "arena_bytes", "arena_used", "arena_peak",
        "queued", "in_flight", "loop_launches", "loop_items", "loop_pauses")]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


# numpy view of dyc_request, for building request arrays in bulk
REQUEST_DTYPE = [("table", "<u4"), ("op", "<u4"), ("keys", "<u8"), ("vals", "<u8"), ("out_vals", "<u8"),
                 ("out_flags", "<u8"), ("n", "<u8")]

_lib = None


def load(require: bool = True):
    """Load libdyc.so; raise if it is missing (there is no fallback path)."""
    global _lib
    if _lib is not None:
        return _lib
    import torch  # noqa: F401  (must precede the CDLL, see module docstring)

    if not os.path.exists(LIB_PATH):
        if not require:
            return None
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m databases_gpu_amd._build` "
            "(or __graft_entry__.build()); the library has no CPU fallback")
    L = ctypes.CDLL(LIB_PATH)
    vp, u64, u32, ci = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int
    sig = {
        "dyc_abi_version": (ci, []),
        "dyc_last_error": (ctypes.c_char_p, []),
        "dyc_config_init": (ci, [ctypes.POINTER(Config)]),
        "dyc_create": (ci, [ctypes.POINTER(Config), u64, ctypes.POINTER(vp)]),
        "dyc_destroy": (ci, [vp]),
        "dyc_insert": (ci, [vp, vp, vp, u64, vp]),
        "dyc_insert_noalloc": (ci, [vp, vp, vp, u64, vp]),
        "dyc_insert_async": (ci, [vp, vp, vp, u64, vp]),
        "dyc_insert_counted": (ci, [vp, vp, vp, u64, vp, vp]),
        "dyc_group_plan_create": (ci, [vp, vp, vp, vp, vp, vp, vp, vp, ctypes.POINTER(vp)]),
        "dyc_group_plan_run": (ci, [vp, vp]),
        "dyc_group_plan_destroy": (ci, [vp]),
        "dyc_wide_group_create": (ci, [vp, u32, vp, ctypes.POINTER(vp)]),
        "dyc_wide_group_destroy": (ci, [vp]),
        "dyc_wide_group_insert": (ci, [vp, vp, vp, vp, vp]),
        "dyc_wide_group_find": (ci, [vp, vp, vp, vp, vp, vp]),
        "dyc_wide_group_erase": (ci, [vp, vp, vp, vp, vp]),
        "dyc_wide_group_table_stats": (ci, [vp, u32, vp]),
        "dyc_wide_group_dump": (ci, [vp, u32, u32, vp, vp, u64, vp]),
        "dyc_wide_group_pending": (ci, [vp, vp, vp, vp, u64, ctypes.POINTER(u64), vp]),
        "dyc_wide_group_clear": (ci, [vp, vp]),
        "dyc_wide_group_insert_per_table": (ci, [vp, vp, vp, vp, vp]),
        "dyc_wide_group_find_per_table": (ci, [vp, vp, vp, vp, vp, vp]),
        "dyc_sync": (ci, [vp]),
        "dyc_find": (ci, [vp, vp, vp, vp, u64, vp]),
        "dyc_erase": (ci, [vp, vp, vp, u64, vp]),
        "dyc_resize_up": (ci, [vp, vp]),
        "dyc_resize_down": (ci, [vp, vp]),
        "dyc_get_stats": (ci, [vp, ctypes.POINTER(Stats)]),
        "dyc_count_occupied": (ci, [vp, ctypes.POINTER(u64), vp]),
        "dyc_dump": (ci, [vp, u32, vp, u64, vp]),
        "dyc_load": (ci, [vp, vp, vp, vp]),
        "dyc_pending": (ci, [vp, vp, u64, ctypes.POINTER(u64), vp]),
        "dyc_clear": (ci, [vp, vp]),
        "dyc_probe_count": (ci, [vp, vp, u64, ctypes.POINTER(u64), ctypes.POINTER(u64), vp]),
        "dyc_upsert": (ci, [vp, vp, vp, u64, vp, vp]),
        "dyc_insert_host": (ci, [vp, vp, vp, u64, vp]),
        "dyc_find_host": (ci, [vp, vp, vp, vp, u64, vp]),
        "dyc_erase_host": (ci, [vp, vp, vp, u64, vp]),
        "dyc_checkpoint_write": (ci, [vp, ctypes.c_char_p, vp]),
        "dyc_checkpoint_read": (ci, [ctypes.c_char_p, ctypes.POINTER(Config), ctypes.POINTER(vp)]),
        "dyc_get_config": (ci, [vp, ctypes.POINTER(Config)]),
        "dyc_set_config": (ci, [vp, ctypes.POINTER(Config)]),
        "dyc_group_create": (ci, [ctypes.POINTER(Config), u32, vp, ctypes.POINTER(vp)]),
        "dyc_group_destroy": (ci, [vp]),
        "dyc_group_get_launch": (ci, [vp, ctypes.POINTER(GroupLaunch)]),
        "dyc_group_set_launch": (ci, [vp, ctypes.POINTER(GroupLaunch)]),
        "dyc_group_insert": (ci, [vp, vp, vp, vp, vp]),
        "dyc_group_find": (ci, [vp, vp, vp, vp, vp, vp]),
        "dyc_group_erase": (ci, [vp, vp, vp, vp, vp]),
        "dyc_group_insert_per_table": (ci, [vp, vp, vp, vp, vp]),
        "dyc_group_find_per_table": (ci, [vp, vp, vp, vp, vp, vp]),
        "dyc_group_table_stats": (ci, [vp, u32, ctypes.POINTER(Stats)]),
        "dyc_group_dump": (ci, [vp, u32, u32, vp, u64, vp]),
        "dyc_group_pending": (ci, [vp, vp, vp, u64, ctypes.POINTER(u64), vp]),
        "dyc_group_clear": (ci, [vp, vp]),
        "dyc_server_config_init": (ci, [ctypes.POINTER(ServerConfig)]),
        "dyc_server_create": (ci, [ctypes.POINTER(Config), ctypes.POINTER(ServerConfig), u32, vp, ctypes.POINTER(vp)]),
        "dyc_server_destroy": (ci, [vp]),
        "dyc_server_submit": (ci, [vp, ctypes.POINTER(Request), ctypes.POINTER(u64)]),
        "dyc_server_poll": (ci, [vp, vp, vp, u32, ctypes.POINTER(u32)]),
        "dyc_server_drain": (ci, [vp]),
        "dyc_server_replay": (ci, [vp, vp, u64, vp, u32, vp, vp]),
        "dyc_server_get_stats": (ci, [vp, ctypes.POINTER(ServerStats)]),
        "dyc_server_table_stats": (ci, [vp, u32, ctypes.POINTER(Stats)]),
        "dyc_server_dump": (ci, [vp, u32, u32, vp, u64]),
        "dyc_wide_create": (ci, [ctypes.POINTER(Config), u64, ctypes.POINTER(vp)]),
        "dyc_wide_destroy": (ci, [vp]),
        "dyc_wide_insert": (ci, [vp, vp, vp, u64, vp]),
        "dyc_wide_upsert": (ci, [vp, vp, vp, u64, vp, vp]),
        "dyc_wide_find": (ci, [vp, vp, vp, vp, u64, vp]),
        "dyc_wide_erase": (ci, [vp, vp, vp, u64, vp]),
        "dyc_wide_resize_up": (ci, [vp, vp]),
        "dyc_wide_resize_down": (ci, [vp, vp]),
        "dyc_wide_get_stats": (ci, [vp, ctypes.POINTER(Stats)]),
        "dyc_wide_clear": (ci, [vp, vp]),
        "dyc_wide_dump": (ci, [vp, u32, vp, vp, u64, vp]),
        "dyc_wide_load": (ci, [vp, vp, vp, vp, vp]),
        "dyc_wide_pending": (ci, [vp, vp, vp, u64, ctypes.POINTER(u64), vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(rc: int, where: str):
    if rc != DYC_OK:
        msg = load().dyc_last_error().decode(errors="replace")
        raise DycError(rc, where, msg)


def default_config(**overrides) -> Config:
    cfg = Config()
    check(load().dyc_config_init(ctypes.byref(cfg)), "dyc_config_init")
    for k, v in overrides.items():
        if not hasattr(cfg, k):
            raise KeyError(f"unknown config field {k}")
        setattr(cfg, k, v)
    return cfg
