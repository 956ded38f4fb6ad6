"""ctypes binding of libp3smi.so (the HIP engine behind include/p3smi.h).

The library is built in-tree by ``build()`` (hipcc, gfx950) and loaded from this package
directory. There is no fallback: if the library is missing or cannot be loaded, every
entry point raises.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_PKG_ROOT = os.path.dirname(_HERE)
_REPO = os.path.dirname(_PKG_ROOT)
LIB_PATH = os.path.join(_HERE, "libp3smi.so")
CSRC = os.path.join(_PKG_ROOT, "csrc")
INCLUDE = os.path.join(_REPO, "include")
SOURCES = ["plan.cpp", "engine.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # no packed-FP32 instructions in device code
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
]


def build(verbose: bool = False) -> str:
    """Compile libp3smi.so for gfx950 in place (always recompiles)."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = LIB_PATH + ".tmp"
    cmd = [hipcc, *HIPCC_FLAGS, f"-I{INCLUDE}", *[os.path.join(CSRC, s) for s in SOURCES], "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed ({r.returncode}) building {LIB_PATH}")
    if verbose:
        sys.stderr.write(r.stderr)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


class P3smiError(RuntimeError):
    pass


class Info(C.Structure):
    _fields_ = [
        ("n", C.c_int32), ("m", C.c_int32), ("npv", C.c_int32), ("npq", C.c_int32),
        ("nnz_y", C.c_int32), ("nnz_j", C.c_int32), ("nnz_lu", C.c_int64), ("n_pairs", C.c_int64),
        ("n_levels", C.c_int32), ("n_blevels", C.c_int32), ("n_slots", C.c_int32), ("ordering", C.c_int32),
        ("rcond0", C.c_double),
        ("lu_g", C.c_int32), ("lu_ns", C.c_int32), ("lu_few_g", C.c_int32), ("lu_few_ns", C.c_int32),
        ("lu_few_max", C.c_int32), ("lu_group", C.c_int32), ("eval_rows", C.c_int32),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("cases", C.c_int32), ("newton_passes", C.c_int32), ("line_search_passes", C.c_int32),
        ("failed_factorizations", C.c_int32), ("guard_trips", C.c_int32), ("reruns", C.c_int32),
        ("single_factorizations", C.c_int64), ("double_factorizations", C.c_int64), ("refinement_steps", C.c_int64),
        ("ms_eval", C.c_double), ("ms_lu", C.c_double), ("ms_step", C.c_double), ("ms_decide", C.c_double),
        ("ms_other", C.c_double), ("ms_refine", C.c_double),
        ("n_eval", C.c_int32), ("n_lu", C.c_int32), ("n_step", C.c_int32), ("n_decide", C.c_int32),
        ("n_refine", C.c_int32), ("pivot_factorizations", C.c_int32),
        ("ms_lu_batch", C.c_double), ("n_lu_batch", C.c_int32), ("lu_batch_cases", C.c_int64),
        ("diagonal_factorizations", C.c_int32),
    ]


# every symbol include/p3smi.h declares
EXPORTS = [
    "p3smi_last_error", "p3smi_version", "p3smi_device_count", "p3smi_device_memory", "p3smi_hbm_probe",
    "p3smi_topology_create", "p3smi_topology_free", "p3smi_topology_info", "p3smi_update_y",
    "p3smi_topology_set_lu_shape", "p3smi_topology_set_eval_tile", "p3smi_topology_set_guard", "p3smi_topology_set_pivoting", "p3smi_dc_start", "p3smi_topology_set_mixed_switch", "p3smi_topology_set_max_chunk", "p3smi_topology_set_mixed",
    "p3smi_host_jacobian", "p3smi_host_factor_solve", "p3smi_host_lu_emulate", "p3smi_host_maps",
    "p3smi_solve_batch", "p3smi_solve_batch_contingency", "p3smi_solve_batch_contingency_stamp",
    "p3smi_outages_branch_flows", "p3smi_outages_voltage_violations",
    "p3smi_batch_create", "p3smi_batch_free", "p3smi_batch_build_sbus", "p3smi_batch_build_sbus_blocks", "p3smi_batch_download_sbus", "p3smi_batch_upload", "p3smi_batch_solve", "p3smi_batch_download",
    "p3smi_topology_release_cache", "p3smi_topology_last_batch", "p3smi_batch_vrange", "p3smi_batch_stats", "p3smi_batch_set_graphs", "p3smi_topology_set_graphs", "p3smi_batch_set_timing", "p3smi_batch_bytes_per_case",
    "p3smi_outages_create", "p3smi_outages_free", "p3smi_outages_islands", "p3smi_outages_islands_noalloc", "p3smi_outages_solve",
    "p3smi_batch_slots_init", "p3smi_batch_load", "p3smi_batch_pass", "p3smi_batch_finished", "p3smi_batch_fetch",
    "p3smi_group_create", "p3smi_group_free", "p3smi_group_solve", "p3smi_group_stats",
]

_lib = None
P = C.c_void_p
_i32 = C.c_int32
_f64 = C.c_double


def lib():
    """The loaded library (loads on first use; raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise P3smiError(f"{LIB_PATH} is not built: run __graft_entry__.build() (hipcc) first")
    L = C.CDLL(LIB_PATH)
    L.p3smi_last_error.restype = C.c_char_p
    L.p3smi_version.restype = C.c_int
    L.p3smi_device_count.argtypes = [P]
    L.p3smi_device_memory.argtypes = [P, P]
    L.p3smi_hbm_probe.argtypes = [_i32, C.c_int64, _i32, P]
    L.p3smi_topology_create.argtypes = [P, P, P, _i32, _i32, P, _i32, P, _i32, _i32, _i32, P]
    L.p3smi_topology_free.argtypes = [P]
    L.p3smi_topology_free.restype = None
    L.p3smi_topology_info.argtypes = [P, P]
    L.p3smi_update_y.argtypes = [P, P, _i32]
    L.p3smi_topology_set_mixed_switch.argtypes = [P, C.c_double]
    L.p3smi_topology_set_lu_shape.argtypes = [P, _i32, _i32, _i32, _i32, _i32, _i32]
    L.p3smi_topology_set_eval_tile.argtypes = [P, _i32]
    L.p3smi_topology_set_guard.argtypes = [P, _f64]
    L.p3smi_topology_set_pivoting.argtypes = [P, _i32]
    L.p3smi_dc_start.argtypes = [P, P, P, P, _i32, P, P]
    L.p3smi_topology_set_mixed.argtypes = [P, _f64, _i32]
    L.p3smi_host_jacobian.argtypes = [P, P, P, P, P, P, P]
    L.p3smi_host_factor_solve.argtypes = [P, P, P, P, P]
    L.p3smi_host_lu_emulate.argtypes = [P, _i32, _i32, _i32, P, P, P, P, P]
    L.p3smi_host_maps.argtypes = [P, P, P, P]
    L.p3smi_solve_batch.argtypes = [P, P, P, _i32, _i32, _i32, _f64, _i32, _i32, P, P, P, P]
    L.p3smi_solve_batch_contingency.argtypes = [P, P, P, P, P, P, _i32, _i32, _f64, _i32, _i32, P, P, P, P]
    L.p3smi_solve_batch_contingency_stamp.argtypes = [P, P, P, P, C.c_int64, P, P, P, _i32, _i32, _f64, _i32, _i32,
                                                      P, P, P, P]
    L.p3smi_batch_create.argtypes = [P, _i32, P]
    L.p3smi_batch_free.argtypes = [P]
    L.p3smi_batch_free.restype = None
    L.p3smi_batch_upload.argtypes = [P, P, P, _i32]
    L.p3smi_batch_solve.argtypes = [P, _i32, _f64, _i32, _i32]
    L.p3smi_batch_download.argtypes = [P, P, P, P, P]
    L.p3smi_outages_branch_flows.argtypes = [P, P, P, P, P, C.c_double, C.c_double, P, P, P, C.c_int32, P, P, P, P,
                                             P]
    L.p3smi_outages_voltage_violations.argtypes = [P, P, P, P, C.c_double, C.c_int32, P, P, P, P, P]
    L.p3smi_batch_build_sbus.argtypes = [P, P, P, C.c_int32, P, P, P, P, C.c_int32, P]
    L.p3smi_batch_build_sbus_blocks.argtypes = [P, P, P, C.c_int32, P, P, P, P, C.c_int32, C.c_int32, P, P]
    L.p3smi_batch_download_sbus.argtypes = [P, P]
    L.p3smi_topology_release_cache.argtypes = [P]
    L.p3smi_topology_last_batch.argtypes = [P, P]
    L.p3smi_topology_set_max_chunk.argtypes = [P, _i32]
    L.p3smi_batch_vrange.argtypes = [P, P, P]
    L.p3smi_batch_stats.argtypes = [P, P]
    L.p3smi_batch_set_graphs.argtypes = [P, C.c_int32]
    L.p3smi_topology_set_graphs.argtypes = [P, C.c_int32]
    L.p3smi_batch_set_timing.argtypes = [P, _i32]
    L.p3smi_batch_bytes_per_case.argtypes = [P, P]
    L.p3smi_outages_create.argtypes = [P, _i32, P, P, P, P, _i32, P, P, _i32, P, P, _i32, P]
    L.p3smi_outages_free.argtypes = [P]
    L.p3smi_outages_free.restype = None
    L.p3smi_outages_islands.argtypes = [P, P, P, P]
    L.p3smi_outages_islands_noalloc.argtypes = [P, P, P, P]
    L.p3smi_outages_solve.argtypes = [P, P, P, _i32, _f64, _i32, _i32, P, P, P, P, P]
    L.p3smi_batch_slots_init.argtypes = [P]
    L.p3smi_batch_load.argtypes = [P, _i32, P, P, P]
    L.p3smi_batch_pass.argtypes = [P, _i32, _f64, _i32, _i32, P]
    L.p3smi_batch_finished.argtypes = [P, P, P]
    L.p3smi_batch_fetch.argtypes = [P, _i32, P, P, P, P, P]
    L.p3smi_group_create.argtypes = [P, _i32, P]
    L.p3smi_group_free.argtypes = [P]
    L.p3smi_group_free.restype = None
    L.p3smi_group_solve.argtypes = [P, _i32, _f64, _i32, _i32, _i32]
    L.p3smi_group_stats.argtypes = [P, P]
    _lib = L
    return L


def check(rc: int, prefix: str = "") -> None:
    if rc != 0:
        msg = lib().p3smi_last_error().decode(errors="replace")
        raise P3smiError(prefix + msg)


def ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def device_count() -> int:
    n = C.c_int32(0)
    lib().p3smi_device_count(C.byref(n))
    return int(n.value)


def device_memory() -> tuple:
    """(free, total) device memory in bytes; raises without a device."""
    f, t = C.c_int64(0), C.c_int64(0)
    check(lib().p3smi_device_memory(C.byref(f), C.byref(t)), "device_memory: ")
    return int(f.value), int(t.value)


def hbm_probe(pattern: int, nbytes: int, repeats: int = 3) -> float:
    """Average milliseconds per launch of an HBM reference pattern over two buffers of
    ``nbytes`` (0: linear copy, 1: 512-byte rows between permuted positions)."""
    ms = C.c_double(0.0)
    check(lib().p3smi_hbm_probe(int(pattern), int(nbytes), int(repeats), C.byref(ms)), "hbm_probe: ")
    return float(ms.value)
