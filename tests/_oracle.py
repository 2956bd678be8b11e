"""ctypes access to the checkers: the C restatement (oracle/libqarp_oracle.so)
and, when built, the reference simulator itself (oracle/_ref/libqarp_ref.so).

Test infrastructure: imported by tests, the smoke check and the CPU-baseline
leg of bench.py only."""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path
from typing import Optional

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
ORACLE_SO = ROOT / "oracle" / "libqarp_oracle.so"
REF_SO = ROOT / "oracle" / "_ref" / "libqarp_ref.so"

_orc = None
_ref = None


def _build_oracle():
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "oracle"], check=True)


def oracle():
    global _orc
    if _orc is None:
        if not ORACLE_SO.exists():
            _build_oracle()
        lib = C.CDLL(str(ORACLE_SO))
        vp, i64, i32 = C.c_void_p, C.c_int64, C.c_int
        lib.orc_last_error.restype = C.c_char_p
        lib.orc_statevector.argtypes = [vp, i32, i32, i32, vp, vp]
        lib.orc_apply_unfused.argtypes = [vp, i32, vp]
        lib.orc_transition.argtypes = [vp, vp, i32, vp, vp]
        lib.orc_apply_observable.argtypes = [vp, i32, vp, vp]
        lib.orc_mt19937.argtypes = [C.c_uint32, i64, vp]
        lib.orc_uniforms.argtypes = [C.c_uint32, i64, vp]
        lib.orc_sample.argtypes = [vp, i32, i64, C.c_uint32, vp]
        lib.orc_sample_uniforms.argtypes = [vp, i32, vp, i64, vp]
        lib.orc_cdf.argtypes = [vp, i32, vp]
        lib.orc_run_gradient.argtypes = [vp, i32, vp, i32, vp, vp]
        lib.orc_run_gradient_phi.argtypes = [vp, i32, vp, i32, vp, vp]
        lib.orc_fuse.argtypes = [vp, i32, i32, i32, C.POINTER(vp)]
        lib.orc_fused_count.argtypes = [vp]
        lib.orc_fused_count.restype = i64
        lib.orc_fused_get.argtypes = [vp, i64, C.POINTER(i32), C.POINTER(i32), vp, vp, i64]
        lib.orc_fused_free.argtypes = [vp]
        lib.orc_gate_matrix.argtypes = [vp, i64, vp, i64]
        _orc = lib
    return _orc


def reference():
    """The reference build, or None where it was not built (GPU box without
    the prebuilt file, or a container without the reference tree)."""
    global _ref
    if _ref is None and REF_SO.exists():
        lib = C.CDLL(str(REF_SO))
        vp, i64, i32 = C.c_void_p, C.c_int64, C.c_int
        lib.ref_last_error.restype = C.c_char_p
        lib.ref_statevector.argtypes = [vp, i32, i32, i32, vp, vp]
        lib.ref_expectation.argtypes = [vp, i32, i32, i32, vp, vp, vp]
        lib.ref_transition.argtypes = [vp, vp, i32, vp, vp]
        lib.ref_fuse.argtypes = [vp, i32, i32, i32, i64, vp, vp, vp, vp, vp, vp]
        lib.ref_sample_state.argtypes = [vp, i32, i32, C.c_uint32, vp]
        lib.ref_run_counts.argtypes = [vp, i32, i32, C.c_uint32, i32, i32, i64, vp, vp, vp]
        lib.ref_mt19937.argtypes = [C.c_uint32, i64, vp]
        lib.ref_uniforms.argtypes = [C.c_uint32, i64, vp]
        lib.ref_discrete.argtypes = [vp, i64, C.c_uint32, i64, vp]
        lib.ref_run_gradient.argtypes = [vp, i32, vp, i32, vp, vp, vp]
        lib.ref_run_gradient_phi.argtypes = [vp, i32, vp, i32, vp, vp, vp]
        _ref = lib
    return _ref


class OracleError(RuntimeError):
    pass


def _check(lib, rc, prefix="orc"):
    if rc != 0:
        msg = getattr(lib, f"{prefix}_last_error")().decode()
        raise OracleError(f"[{rc}] {msg}")


def _cptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data


def statevector(packed, n, fusion_max=3, fusion_min=12, init=None, lib="orc"):
    out = np.zeros(1 << n, dtype=np.complex128)
    init = None if init is None else np.ascontiguousarray(init, dtype=np.complex128)
    L = oracle() if lib == "orc" else reference()
    fn = L.orc_statevector if lib == "orc" else L.ref_statevector
    _check(L, fn(packed.ref, n, fusion_max, fusion_min, _cptr(init), out.ctypes.data), lib)
    return out


def transition(bra, ket, n, pobs, lib="orc"):
    out = np.zeros(2)
    L = oracle() if lib == "orc" else reference()
    fn = L.orc_transition if lib == "orc" else L.ref_transition
    bra = np.ascontiguousarray(bra, dtype=np.complex128)
    ket = np.ascontiguousarray(ket, dtype=np.complex128)
    _check(L, fn(bra.ctypes.data, ket.ctypes.data, n, pobs.ref, out.ctypes.data), lib)
    return complex(out[0], out[1])


def apply_observable(psi, n, pobs):
    out = np.zeros(1 << n, dtype=np.complex128)
    psi = np.ascontiguousarray(psi, dtype=np.complex128)
    _check(oracle(), oracle().orc_apply_observable(psi.ctypes.data, n, pobs.ref, out.ctypes.data))
    return out


def mt19937(seed, n, lib="orc"):
    out = np.zeros(n, dtype=np.uint32)
    if lib == "orc":
        oracle().orc_mt19937(seed, n, out.ctypes.data)
    else:
        reference().ref_mt19937(seed, n, out.ctypes.data)
    return out


def uniforms(seed, n, lib="orc"):
    out = np.zeros(n, dtype=np.float64)
    if lib == "orc":
        oracle().orc_uniforms(seed, n, out.ctypes.data)
    else:
        reference().ref_uniforms(seed, n, out.ctypes.data)
    return out


def sample(sv, n, n_shots, seed, lib="orc"):
    out = np.zeros(n_shots, dtype=np.uint64)
    sv = np.ascontiguousarray(sv, dtype=np.complex128)
    if lib == "orc":
        _check(oracle(), oracle().orc_sample(sv.ctypes.data, n, n_shots, seed, out.ctypes.data))
    else:
        _check(reference(), reference().ref_sample_state(sv.ctypes.data, n, n_shots, seed,
                                                         out.ctypes.data), "ref")
    return out


def sample_uniforms(sv, n, u):
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = np.zeros(u.size, dtype=np.uint64)
    sv = np.ascontiguousarray(sv, dtype=np.complex128)
    _check(oracle(), oracle().orc_sample_uniforms(sv.ctypes.data, n, u.ctypes.data, u.size,
                                                  out.ctypes.data))
    return out


def cdf(sv, n):
    out = np.zeros(1 << n, dtype=np.float64)
    sv = np.ascontiguousarray(sv, dtype=np.complex128)
    _check(oracle(), oracle().orc_cdf(sv.ctypes.data, n, out.ctypes.data))
    return out


def run_gradient(packed, n, pobs, n_params, values=None, init=None, lib="orc"):
    grad = np.zeros(n_params)
    init = None if init is None else np.ascontiguousarray(init, dtype=np.complex128)
    if lib == "orc":
        _check(oracle(), oracle().orc_run_gradient(packed.ref, n, pobs.ref, n_params, _cptr(init),
                                                   grad.ctypes.data))
    else:
        v = np.ascontiguousarray(values, dtype=np.float64)
        _check(reference(), reference().ref_run_gradient(packed.ref, n, pobs.ref, n_params,
                                                         v.ctypes.data, _cptr(init),
                                                         grad.ctypes.data), "ref")
    return grad


def run_gradient_phi(packed, n, phi, n_params, values=None, init=None, lib="orc"):
    grad = np.zeros(n_params)
    phi = np.ascontiguousarray(phi, dtype=np.complex128)
    init = None if init is None else np.ascontiguousarray(init, dtype=np.complex128)
    if lib == "orc":
        _check(oracle(), oracle().orc_run_gradient_phi(packed.ref, n, phi.ctypes.data, n_params,
                                                       _cptr(init), grad.ctypes.data))
    else:
        v = np.ascontiguousarray(values, dtype=np.float64)
        _check(reference(), reference().ref_run_gradient_phi(packed.ref, n, phi.ctypes.data,
                                                             n_params, v.ctypes.data, _cptr(init),
                                                             grad.ctypes.data), "ref")
    return grad


def fuse(packed, n, fusion_max, fusion_min):
    """Oracle fusion: list of (src, qubits, matrix-or-None)."""
    L = oracle()
    h = C.c_void_p()
    _check(L, L.orc_fuse(packed.ref, n, fusion_max, fusion_min, C.byref(h)))
    out = []
    try:
        for i in range(L.orc_fused_count(h)):
            src, k = C.c_int(), C.c_int()
            qs = np.zeros(64, dtype=np.uint32)
            m = np.zeros((256, 256), dtype=np.complex128)
            L.orc_fused_get(h, i, C.byref(src), C.byref(k), qs.ctypes.data, m.ctypes.data, 65536)
            if src.value >= 0:
                out.append((src.value, None, None))
            else:
                d = 1 << k.value
                out.append((-1, tuple(int(q) for q in qs[: k.value]),
                            m.ravel()[: d * d].reshape(d, d).copy()))
    finally:
        L.orc_fused_free(h)
    return out


def ref_fuse(packed, n, fusion_max, fusion_min, cap=100000):
    """Reference fusion: list of (gate, qubits, params, matrix-or-None)."""
    L = reference()
    n_out = C.c_int64()
    gates = np.zeros(cap, np.int32)
    nqs = np.zeros(cap, np.int32)
    qubits = np.zeros((cap, 8), np.uint32)
    params = np.zeros((cap, 4))
    mats = np.zeros((cap, 1024), np.complex128)
    _check(L, L.ref_fuse(packed.ref, n, fusion_max, fusion_min, cap, C.byref(n_out),
                         gates.ctypes.data, nqs.ctypes.data, qubits.ctypes.data,
                         params.ctypes.data, mats.ctypes.data), "ref")
    out = []
    for i in range(n_out.value):
        k = int(nqs[i])
        m = None
        if gates[i] == 46:
            d = 1 << k
            m = mats[i, : d * d].reshape(d, d).copy()
        out.append((int(gates[i]), tuple(int(q) for q in qubits[i, : min(k, 8)]),
                    tuple(params[i]), m))
    return out
