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
        lib.orc_permute.argtypes = [vp, i32, vp, i32, vp, vp]
        lib.orc_controlled_powers.argtypes = [vp, i32, vp, vp, i32, vp, i32, vp]
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
        lib.ref_program_statevector.argtypes = [vp, i32, i32, i32, i32, vp, vp]
        lib.ref_run_trajectory.argtypes = [vp, vp, vp, vp, i32, i32, C.c_uint32, i64, vp, vp, vp, vp, vp, i64]
        lib.ref_statevector_fingerprint.argtypes = [vp, i32, i32, i32, vp, i64, vp, vp, vp]
        lib.ref_run_noisy.argtypes = [vp, vp, vp, vp, vp, i32, i32, i32, i32, C.c_uint32, i64, vp, vp, vp, vp, vp,
                                      i64]
        lib.ref_uniform_int.argtypes = [C.c_uint32, C.c_uint64, C.c_uint64, i64, vp]
        lib.ref_permutation_table.argtypes = [vp, vp, i32, C.c_uint32, C.c_uint32, vp, vp]
        lib.ref_local_unitary_of.argtypes = [vp, vp, i32, vp]
        lib.ref_commands_digest.argtypes = [vp]
        lib.ref_commands_digest.restype = C.c_uint64
        lib.ref_local_commands_digest.argtypes = [vp, vp, i32]
        lib.ref_local_commands_digest.restype = C.c_uint64
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


def fingerprint_indices(n, count, seed):
    """Seeded distinct amplitude indices (ascending) for a state fingerprint;
    index 0 is always included."""
    rng = np.random.default_rng(seed)
    dim = 1 << n
    if count >= dim:
        return np.arange(dim, dtype=np.uint64)
    idx = np.unique(rng.integers(0, dim, size=count, dtype=np.uint64))
    return np.union1d(idx, np.array([0], np.uint64)).astype(np.uint64)


def ref_fingerprint(packed, n, fusion_max, fusion_min, idx, pobs=None):
    """Reference statevector reduced to (amplitudes at idx, norm, <Z_q> for
    every q, <obs> or None) without a host copy of the state."""
    idx = np.ascontiguousarray(idx, dtype=np.uint64)
    amps = np.zeros(idx.size, dtype=np.complex128)
    stats = np.zeros(n + 2)
    L = reference()
    _check(L, L.ref_statevector_fingerprint(packed.ref, n, fusion_max, fusion_min, idx.ctypes.data,
                                            idx.size, amps.ctypes.data,
                                            None if pobs is None else pobs.ref, stats.ctypes.data), "ref")
    return amps, float(stats[0]), stats[1:n + 1].copy(), (None if pobs is None else float(stats[n + 1]))


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


# ── structured programs ────────────────────────────────────────────────────

def permute(psi, n, qubits, table):
    """apply_permutation restated (orc_permute)."""
    inp = np.ascontiguousarray(psi, dtype=np.complex128)
    out = np.zeros(1 << n, dtype=np.complex128)
    qs = np.ascontiguousarray(qubits, dtype=np.uint32)
    t = np.ascontiguousarray(table, dtype=np.uint64)
    L = oracle()
    _check(L, L.orc_permute(inp.ctypes.data, n, qs.ctypes.data, qs.size, t.ctypes.data, out.ctypes.data))
    return out


def controlled_powers(psi, n, controls, exponents, targets, matrix):
    """apply_controlled_powers restated (orc_controlled_powers)."""
    st = np.array(psi, dtype=np.complex128)
    cs = np.ascontiguousarray(controls, dtype=np.uint32)
    es = np.ascontiguousarray(exponents, dtype=np.uint64)
    ts = np.ascontiguousarray(targets, dtype=np.uint32)
    m = np.ascontiguousarray(matrix, dtype=np.complex128)
    L = oracle()
    _check(L, L.orc_controlled_powers(st.ctypes.data, n, cs.ctypes.data, es.ctypes.data, cs.size, ts.ctypes.data,
                                      ts.size, m.ctypes.data))
    return st


def program_statevector(kernels, n, init=None, fusion_max=3, fusion_min=12, lib="orc"):
    """A structured program's statevector: the oracle composes its kernels
    in order (gates: orc_statevector from the current state; dense: the
    matrix as one Custom op; permutation / controlled powers: the restated
    kernels); lib="ref" runs QarpSimulator::program_statevector of the
    reference build.  kernels: the tuples of openqarp_amd.program.Program."""
    from openqarp_amd._abi import pack_stream
    from openqarp_amd.ir import Command, GateType
    psi = np.zeros(1 << n, dtype=np.complex128)
    if init is None:
        psi[0] = 1.0
    else:
        psi[:] = init
    if lib == "orc":
        for k in kernels:
            if k[0] == "gates":
                psi = statevector(pack_stream(k[1]), n, fusion_max, fusion_min, init=psi)
            elif k[0] == "dense":
                psi = statevector(pack_stream([Command(GateType.Custom, qubits=k[1], unitary=k[2])]), n,
                                  fusion_max, fusion_min, init=psi)
            elif k[0] == "permutation":
                psi = permute(psi, n, k[1], k[2])
            else:
                psi = controlled_powers(psi, n, k[1], k[2], k[3], k[4])
        return psi

    class K(C.Structure):
        _fields_ = [("kind", C.c_int32), ("k", C.c_int32), ("n_controls", C.c_int32), ("unused", C.c_int32),
                    ("qubits", C.c_void_p), ("table", C.c_void_p), ("matrix", C.c_void_p),
                    ("controls", C.c_void_p), ("exponents", C.c_void_p), ("gates", C.c_void_p)]
    keep = []
    arr = (K * max(len(kernels), 1))()

    def hold(a):
        keep.append(a)
        return a.ctypes.data
    for i, k in enumerate(kernels):
        e = arr[i]
        if k[0] == "gates":
            ps = pack_stream(k[1])
            keep.append(ps)
            e.kind, e.gates = 0, C.addressof(ps.struct)
        elif k[0] == "permutation":
            e.kind, e.k = 1, len(k[1])
            e.qubits = hold(np.ascontiguousarray(k[1], dtype=np.uint32))
            e.table = hold(np.ascontiguousarray(k[2], dtype=np.uint64))
        elif k[0] == "dense":
            e.kind, e.k = 2, len(k[1])
            e.qubits = hold(np.ascontiguousarray(k[1], dtype=np.uint32))
            e.matrix = hold(np.ascontiguousarray(k[2], dtype=np.complex128))
        else:
            e.kind, e.k, e.n_controls = 3, len(k[3]), len(k[1])
            e.controls = hold(np.ascontiguousarray(k[1], dtype=np.uint32))
            e.exponents = hold(np.ascontiguousarray(k[2], dtype=np.uint64))
            e.qubits = hold(np.ascontiguousarray(k[3], dtype=np.uint32))
            e.matrix = hold(np.ascontiguousarray(k[4], dtype=np.complex128))
    out = np.zeros(1 << n, dtype=np.complex128)
    L = reference()
    init_c = np.ascontiguousarray(psi)
    _check(L, L.ref_program_statevector(C.cast(arr, C.c_void_p), len(kernels), n, fusion_max, fusion_min,
                                        init_c.ctypes.data, out.ctypes.data), "ref")
    return out


# ── trajectories (QarpSimulator::run with mid-circuit measurement) ───────────

# ── noise (the trajectory path) ─────────────────────────────────────────────
# A model is described as builder calls [(kind, (p0, p1, p2), [gate ids])]:
# kind 0 make_depolarizing(p0), 1 make_pauli(p0, p1, p2), 2 make_bit_flip(p0),
# 3 make_amplitude_damping(p0), composed left to right (the right one wins
# per gate), as the reference's operator+ does.

class _RefNoiseEntry(C.Structure):
    _fields_ = [("kind", C.c_int32), ("n_gates", C.c_int32), ("p", C.c_double * 3), ("gates", C.c_void_p)]


def _two_qubit(g):
    from openqarp_amd.ir import GateType as G
    return G(g) in {G.CX, G.CY, G.CZ, G.SWAP, G.ECR, G.iSWAP, G.iSWAPdg, G.CH, G.CS, G.CSdg, G.CSX, G.CSXdg,
                    G.CRx, G.CRy, G.CRz, G.CP, G.RZZ, G.RXX, G.RYY, G.CU}


def noise_channels(entries):
    """{gate id: channel} of a described model, restating the builders
    (device/noise_model.cpp:221-299): ("pauli", px, py, pz), ("pauli2", [15]),
    ("kq", p), ("kraus", [K0, K1])."""
    from openqarp_amd.ir import GateType as G
    out = {}
    for kind, p, gates in entries:
        for g in gates:
            g = int(g)
            if kind == 0:
                if g == int(G.MCZ):
                    out[g] = ("kq", p[0])
                elif _two_qubit(g):
                    out[g] = ("pauli2", [p[0] / 15.0] * 15)
                else:
                    out[g] = ("pauli", p[0] / 3.0, p[0] / 3.0, p[0] / 3.0)
            elif kind in (1, 2):
                out[g] = ("pauli", p[0], p[1] if kind == 1 else 0.0, p[2] if kind == 1 else 0.0)
            else:
                k0 = np.array([[1.0, 0.0], [0.0, np.sqrt(1.0 - p[0])]], dtype=np.complex128)
                k1 = np.array([[0.0, np.sqrt(p[0])], [0.0, 0.0]], dtype=np.complex128)
                out[g] = ("kraus", [k0, k1])
    return out


class _Words:
    """One shot's std::mt19937 stream (oracle words), read as uniforms
    (generate_canonical from two words) and as uniform_int_distribution's
    integers (Lemire's multiply-and-reject on 32-bit words)."""

    def __init__(self, seed, n):
        self.w = mt19937(seed, n).astype(np.uint64)
        self.i = 0

    def word(self):
        v = int(self.w[self.i])
        self.i += 1
        return v

    def uniform(self):
        w0, w1 = self.word(), self.word()
        u = (float(w0) + float(w1) * 4294967296.0) / 18446744073709551616.0
        return u if u < 1.0 else float(np.nextafter(1.0, 0.0))

    def below(self, r):
        if r < 0xFFFFFFFF:
            m = r + 1
            prod = self.word() * m
            if (prod & 0xFFFFFFFF) < m:
                t = ((1 << 32) - m) % m
                while (prod & 0xFFFFFFFF) < t:
                    prod = self.word() * m
            return prod >> 32
        if r == 0xFFFFFFFF:
            return self.word()
        while True:
            tmp = (1 << 32) * self.below(r >> 32)
            ret = tmp + self.word()
            if tmp <= ret <= r:
                return ret


def uniform_ints(seed, lo, hi, count, lib="orc"):
    """count draws of std::uniform_int_distribution<uint64_t>(lo, hi) over
    std::mt19937(seed)."""
    if lib == "ref":
        out = np.zeros(count, np.uint64)
        reference().ref_uniform_int(seed, lo, hi, count, out.ctypes.data)
        return out
    w = _Words(seed, 4 * count + 64)
    return np.array([lo + w.below(hi - lo) for _ in range(count)], dtype=np.uint64)


def _pauli_1q(psi, q, letter):
    idx = np.arange(psi.size)
    bit = ((idx >> q) & 1).astype(bool)
    if letter == 1:
        return psi[idx ^ (1 << q)]
    if letter == 2:
        src = psi[idx ^ (1 << q)]
        return np.where(bit, 1j * src, -1j * src)
    if letter == 3:
        return np.where(bit, -psi, psi)
    return psi


def _apply_channel(ch, qs, psi, words):
    """apply_channel restated (device/noise_model.cpp:86-186)."""
    if ch[0] == "pauli":
        r = words.uniform()
        acc = ch[1]
        if r < acc:
            return _pauli_1q(psi, qs[0], 1)
        acc += ch[2]
        if r < acc:
            return _pauli_1q(psi, qs[0], 2)
        acc += ch[3]
        if r < acc:
            return _pauli_1q(psi, qs[0], 3)
        return psi
    if ch[0] == "pauli2":
        r = words.uniform()
        acc = 0.0
        for i, p in enumerate(ch[1]):
            acc += p
            if r < acc:
                packed = i + 1
                psi = _pauli_1q(psi, qs[0], packed & 3)
                return _pauli_1q(psi, qs[1], (packed >> 2) & 3)
        return psi
    if ch[0] == "kq":
        k = len(qs)
        if k == 0 or k > 31:
            return psi
        if words.uniform() >= ch[1]:
            return psi
        code = 1 + words.below((1 << (2 * k)) - 2)
        for i in range(k):
            psi = _pauli_1q(psi, qs[i], (code >> (2 * i)) & 3)
        return psi
    q = qs[0]
    idx = np.arange(psi.size)
    zero = ((idx >> q) & 1) == 0
    a, b = psi[zero], psi[idx[zero] | (1 << q)]
    r00 = np.add.accumulate((np.conj(a) * a).real)[-1]
    r11 = np.add.accumulate((np.conj(b) * b).real)[-1]
    r01 = np.add.accumulate(a * np.conj(b))[-1]
    rho = np.array([[r00, r01], [np.conj(r01), r11]], dtype=np.complex128)
    probs = [max(0.0, float(np.trace(K @ rho @ K.conj().T).real)) for K in ch[1]]
    total = 0.0
    for p in probs:
        total += p
    if total <= 0.0:
        return psi
    probs = [p / total for p in probs]
    r = words.uniform()
    acc = 0.0
    for i, p in enumerate(probs):
        acc += p
        if r < acc:
            pu = p * total
            if pu <= 0.0:
                return psi
            m = ch[1][i] * (1.0 / np.sqrt(pu))
            out = psi.copy()
            out[zero] = m[0, 0] * a + m[0, 1] * b
            out[idx[zero] | (1 << q)] = m[1, 0] * a + m[1, 1] * b
            return out
    return psi


def run_trajectory(commands, n, n_shots, seed, lib="orc", noise=None, enabled=True):
    """(counts dict, cbit history list) of QarpSimulator::run.  lib="ref":
    the reference build.  lib="orc": a per-shot restatement of the
    reference's trajectory loop (qarp_simulator.cpp:1038-1211,
    apply_command_trajectory :600-689): the unitary prefix once, then per
    shot std::mt19937(seed + shot) (orc_uniforms), each Measure / Reset draws
    u, outcome = u >= P(0) (sum of |a|^2 in index order), projection and
    scaling by sqrt(1 / p), conditions and branches on the shot's register,
    and the final inline CDF walk."""
    from openqarp_amd._abi import pack_stream
    from openqarp_amd.ir import GateType, cbit_register_width, Command
    from openqarp_amd.program import needs_trajectory
    if lib == "ref":
        L = reference()
        cap = 1 << n
        keys = np.zeros(cap, np.uint64)
        cnt = np.zeros(cap, np.int32)
        nk = C.c_int64()
        nc = C.c_int32()
        width = cbit_register_width(commands)
        hist = np.zeros(max(1, n_shots * width), np.uint8)
        ps = pack_stream(commands)
        ccount = np.array([len(c.condition_bits) for c in commands], dtype=np.int32)
        cbits = np.array([b for c in commands for b in c.condition_bits] or [0], dtype=np.uint32)
        cvals = np.array([bool(v) for c in commands for v in c.condition_values] or [0], dtype=np.uint8)
        if noise is None:
            _check(L, L.ref_run_trajectory(ps.ref, ccount.ctypes.data, cbits.ctypes.data, cvals.ctypes.data, n,
                                           n_shots, seed, cap, C.byref(nk), keys.ctypes.data, cnt.ctypes.data,
                                           C.byref(nc), hist.ctypes.data, hist.size), "ref")
        else:
            gate_arrays = [np.ascontiguousarray(g, dtype=np.int32) for _, _, g in noise]
            ents = (_RefNoiseEntry * max(1, len(noise)))()
            for e, ((kind, p, _), ga) in enumerate(zip(noise, gate_arrays)):
                ents[e].kind, ents[e].n_gates = kind, ga.size
                for k in range(3):
                    ents[e].p[k] = float(p[k]) if k < len(p) else 0.0
                ents[e].gates = ga.ctypes.data
            _check(L, L.ref_run_noisy(ps.ref, ccount.ctypes.data, cbits.ctypes.data, cvals.ctypes.data, ents,
                                      len(noise), 1 if enabled else 0, n, n_shots, seed, cap, C.byref(nk),
                                      keys.ctypes.data, cnt.ctypes.data, C.byref(nc), hist.ctypes.data,
                                      hist.size), "ref")
        counts = {int(keys[i]): int(cnt[i]) for i in range(nk.value)}
        h = hist[: n_shots * nc.value].reshape(n_shots, nc.value).astype(bool)
        return counts, [list(map(bool, r)) for r in h]
    cmds = list(commands)
    channels = noise_channels(noise) if (noise and enabled) else {}
    split = 0 if channels else next((i for i, c in enumerate(cmds) if needs_trajectory(c)), len(cmds))
    base = statevector(pack_stream(cmds[:split]), n, 3, 12)
    suffix = cmds[split:]
    width = cbit_register_width(cmds)
    dim = 1 << n
    ends = {}
    stack = []
    for i, c in enumerate(suffix):
        if c.gate == GateType.BranchBegin:
            stack.append([i, None])
        elif c.gate == GateType.BranchElse:
            stack[-1][1] = i
        elif c.gate == GateType.BranchEnd:
            b, e = stack.pop()
            ends[b] = (e, i)
    counts, hist = {}, []
    noiseless = {GateType.Measure, GateType.Reset, GateType.Barrier, GateType.GPhase}
    for shot in range(n_shots):
        words = _Words((seed + shot) & 0xFFFFFFFF, 8 * len(suffix) + 256)
        reg = [False] * width
        psi = base.copy()

        def draw():
            return words.uniform()

        def holds(c):
            return all(reg[b] == bool(v) for b, v in zip(c.condition_bits, c.condition_values))

        def run(lo, hi):
            nonlocal psi
            i = lo
            while i < hi:
                c = suffix[i]
                if c.gate == GateType.BranchBegin:
                    e, end = ends[i]
                    if holds(c):
                        run(i + 1, e if e is not None else end)
                    elif e is not None:
                        run(e + 1, end)
                    i = end + 1
                    continue
                i += 1
                if c.condition_bits and not holds(c):
                    continue
                if c.gate in (GateType.Measure, GateType.Reset):
                    q = c.qubits[0]
                    idx = np.arange(dim)
                    zero = ((idx >> q) & 1) == 0
                    pz = psi[zero]
                    p0 = float(np.add.accumulate(pz.real * pz.real + pz.imag * pz.imag)[-1])
                    one = draw() >= p0
                    p = max(1.0 - p0, 0.0) if one else max(p0, 0.0)
                    f = np.sqrt(1.0 / p)
                    psi = np.where(zero != one, psi * f, 0)
                    if c.gate == GateType.Reset and one:
                        psi = psi[idx ^ (1 << q)]
                    elif c.gate == GateType.Measure and c.cbits:
                        reg[c.cbits[0]] = bool(one)
                elif c.gate != GateType.Barrier:
                    bare = Command(c.gate, qubits=c.qubits, params=c.params, unitary=c.unitary)
                    psi = statevector(pack_stream([bare]), n, 0, 0, init=psi)
                if c.gate not in noiseless and int(c.gate) in channels:
                    psi = _apply_channel(channels[int(c.gate)], list(c.qubits), psi, words)
        run(0, len(suffix))
        probs = psi.real * psi.real + psi.imag * psi.imag
        run = np.add.accumulate(probs)             # the sequential running sum from 0.0
        threshold = draw() * run[-1]
        out = min(int(np.searchsorted(run, threshold, side="left")), dim - 1)
        counts[out] = counts.get(out, 0) + 1
        hist.append(list(reg))
    return counts, hist


def ref_permutation_table(packed, qubits, max_rest, max_work_log2):
    """permutation_table of the reference build (int64 table or None)."""
    qs = np.ascontiguousarray(qubits, dtype=np.uint32)
    out = np.zeros(1 << qs.size, dtype=np.uint64)
    found = C.c_int32(0)
    L = reference()
    _check(L, L.ref_permutation_table(packed.ref, qs.ctypes.data, qs.size, max_rest, max_work_log2,
                                      C.byref(found), out.ctypes.data), "ref")
    return out.astype(np.int64) if found.value else None


def ref_local_unitary_of(packed, qubits):
    qs = np.ascontiguousarray(qubits, dtype=np.uint32)
    d = 1 << qs.size
    out = np.zeros(d * d, dtype=np.complex128)
    L = reference()
    _check(L, L.ref_local_unitary_of(packed.ref, qs.ctypes.data, qs.size, out.ctypes.data), "ref")
    return out.reshape(d, d)


def ref_commands_digest(packed):
    return int(reference().ref_commands_digest(packed.ref))


def ref_local_commands_digest(packed, qubits):
    qs = np.ascontiguousarray(qubits, dtype=np.uint32)
    return int(reference().ref_local_commands_digest(packed.ref, qs.ctypes.data, qs.size))
