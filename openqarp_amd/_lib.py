"""Loads libqarp_hip.so (built in-tree by openqarp_amd.build) and declares
the C signatures of include/qarp_hip.h.

There is no fallback: without the library, or without a visible GPU, every
entry point raises BackendUnavailable."""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

from .errors import BackendUnavailable, CapabilityError

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libqarp_hip.so"

_lib = None

EXPORTS = [
    "qhip_last_error", "qhip_device_count", "qhip_synchronize",
    "qhip_state_create", "qhip_state_destroy", "qhip_state_info", "qhip_state_reset",
    "qhip_state_upload", "qhip_state_download", "qhip_state_upload_c128",
    "qhip_state_download_c128", "qhip_state_copy", "qhip_state_norm2", "qhip_state_device_ptr",
    "qhip_fuse", "qhip_fused_count", "qhip_fused_get", "qhip_fused_destroy",
    "qhip_schedule_inspect", "qhip_fused_pass",
    "qhip_plan_create", "qhip_plan_apply", "qhip_plan_apply_timed", "qhip_plan_apply_graph",
    "qhip_plan_update", "qhip_plan_destroy",
    "qhip_apply", "qhip_apply_terminal", "qhip_state_permute", "qhip_state_controlled_powers",
    "qhip_state_set_basis", "qhip_compact", "qhip_shots_broadcast", "qhip_shots_restore", "qhip_shots_prob0",
    "qhip_shots_project", "qhip_shots_apply1q", "qhip_shots_rho1", "qhip_shots_sample", "qhip_mt_uniforms",
    "qhip_transition", "qhip_transition_noalloc", "qhip_apply_observable",
    "qhip_sample", "qhip_sample_uniforms", "qhip_probabilities", "qhip_exact",
    "qhip_run_gradient", "qhip_run_gradient_phi", "qhip_batch_energy_grad",
    "qhip_batch_create", "qhip_batch_run", "qhip_batch_update", "qhip_batch_destroy", "qhip_batch_stats",
    "qhip_permutation_table", "qhip_fnv1a_words",
]


def _declare(lib):
    vp, i32, i64, u32, u64, dp = C.c_void_p, C.c_int32, C.c_int64, C.c_uint32, C.c_uint64, C.c_void_p
    sig = {
        "qhip_last_error": [C.c_char_p, C.c_size_t],
        "qhip_device_count": [C.POINTER(i32)],
        "qhip_synchronize": [],
        "qhip_state_create": [i32, i32, C.POINTER(vp)],
        "qhip_state_destroy": [vp],
        "qhip_state_info": [vp, C.POINTER(i32), C.POINTER(i32)],
        "qhip_state_reset": [vp],
        "qhip_state_upload": [vp, vp, u64],
        "qhip_state_download": [vp, vp, u64],
        "qhip_state_upload_c128": [vp, vp, u64],
        "qhip_state_download_c128": [vp, vp, u64],
        "qhip_state_copy": [vp, vp],
        "qhip_state_norm2": [vp, C.POINTER(C.c_double)],
        "qhip_state_device_ptr": [vp, C.POINTER(vp)],
        "qhip_fuse": [vp, vp, i32, C.POINTER(vp)],
        "qhip_fused_count": [vp, C.POINTER(i64)],
        "qhip_fused_get": [vp, i64, C.POINTER(i32), C.POINTER(i32), vp, vp, vp, i64],
        "qhip_fused_destroy": [vp],
        "qhip_schedule_inspect": [vp, vp, i32, i32, C.POINTER(vp)],
        "qhip_fused_pass": [vp, i64, C.POINTER(i32)],
        "qhip_plan_create": [vp, vp, i32, i32, C.POINTER(vp)],
        "qhip_plan_apply": [vp, vp, vp],
        "qhip_plan_apply_timed": [vp, vp, vp, i64],
        "qhip_plan_apply_graph": [vp, vp, vp],
        "qhip_plan_update": [vp, vp, vp, C.POINTER(i32)],
        "qhip_plan_destroy": [vp],
        "qhip_apply": [vp, vp, vp, vp],
        "qhip_apply_terminal": [vp, vp, vp, vp],
        "qhip_state_permute": [vp, vp, i32, vp],
        "qhip_state_set_basis": [vp, vp, i64, u64],
        "qhip_compact": [vp, u64, C.c_double, vp, vp, i64, C.POINTER(i64), C.POINTER(C.c_double)],
        "qhip_shots_broadcast": [vp, vp],
        "qhip_shots_restore": [vp, vp, vp, u64],
        "qhip_shots_prob0": [vp, i32, u64, vp],
        "qhip_shots_project": [vp, i32, u64, vp, vp],
        "qhip_shots_sample": [vp, u64, vp, i64, vp],
        "qhip_shots_apply1q": [vp, i32, u64, vp, vp],
        "qhip_shots_rho1": [vp, i32, u64, vp],
        "qhip_mt_uniforms": [u32, i64, vp],
        "qhip_state_controlled_powers": [vp, vp, vp, vp, i32, vp, i32, vp],
        "qhip_transition": [vp, vp, vp, vp],
        "qhip_transition_noalloc": [vp, vp, vp, vp],
        "qhip_apply_observable": [vp, vp, vp],
        "qhip_sample": [vp, i64, u32, vp],
        "qhip_sample_uniforms": [vp, vp, i64, vp],
        "qhip_probabilities": [vp, vp, u64],
        "qhip_exact": [vp, C.c_double, vp, vp, i64, C.POINTER(i64)],
        "qhip_run_gradient": [vp, vp, i32, vp, vp, vp, vp],
        "qhip_run_gradient_phi": [vp, vp, i32, vp, vp, vp],
        "qhip_batch_energy_grad": [vp, i32, vp, vp, vp, vp, i32, vp, vp, vp],
        "qhip_batch_create": [vp, i32, vp, vp, vp, vp, i32, i32, C.POINTER(vp)],
        "qhip_batch_run": [vp, vp, vp, vp, i32],
        "qhip_batch_update": [vp, vp, C.POINTER(i32)],
        "qhip_batch_destroy": [vp],
        "qhip_batch_stats": [vp, vp],
        "qhip_permutation_table": [vp, i32, u32, u32, vp, vp],
        "qhip_fnv1a_words": [vp, i64, u64, vp],
    }
    for name, args in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = C.c_int


def load():
    """The loaded library (raises BackendUnavailable when it is not built)."""
    global _lib
    if _lib is None:
        path = Path(os.environ.get("QHIP_LIBRARY", LIB_PATH))
        if not path.exists():
            raise BackendUnavailable(
                f"{path} is not built: run `python -m openqarp_amd.build` (hipcc, gfx950)")
        lib = C.CDLL(str(path))
        _declare(lib)
        _lib = lib
    return _lib


def library_path() -> Path:
    return Path(os.environ.get("QHIP_LIBRARY", LIB_PATH))


def check(rc: int) -> None:
    """Map a status code onto the reference's exception types."""
    if rc == 0:
        return
    buf = C.create_string_buffer(4096)
    load().qhip_last_error(buf, len(buf))
    msg = buf.value.decode(errors="replace")
    if rc == 1:
        raise ValueError(msg)
    if rc == 3:
        raise CapabilityError(msg)
    if rc == 4:
        raise BackendUnavailable(msg) if "no HIP device" in msg else RuntimeError(msg)
    raise RuntimeError(msg)


def device_count() -> int:
    n = C.c_int32(0)
    load().qhip_device_count(C.byref(n))
    return n.value
