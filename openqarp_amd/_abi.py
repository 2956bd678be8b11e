"""ctypes image of include/qarp_hip.h and the packing of Python IR into it.

The stream and observable layouts are shared with every C consumer of the
header; ``pack_stream`` / ``pack_observable`` return objects that keep the
numpy buffers alive for as long as the ctypes structs point at them.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Mapping, Optional, Sequence

import numpy as np

from .ir import Command, GateType, affine_coefficients
from .errors import CapabilityError

QHIP_OK, QHIP_INVALID, QHIP_RUNTIME, QHIP_CAPABILITY, QHIP_DEVICE, QHIP_NOMEM = range(6)
QHIP_C128, QHIP_C64 = 0, 1
QHIP_OP_CONDITIONAL = 1

OP_DTYPE = np.dtype({
    "names": ["gate", "n_qubits", "qubit_offset", "n_params", "flags", "params",
              "unitary_offset", "dep_offset", "n_deps", "cbit"],
    "formats": [np.int32, np.int32, np.int64, np.int32, np.int32, (np.float64, 4),
                np.int64, np.int64, np.int32, np.int32],
    "offsets": [0, 4, 8, 16, 20, 24, 56, 64, 72, 76],
    "itemsize": 80,
})

TERM_DTYPE = np.dtype({
    "names": ["offset", "n_letters", "reserved", "coeff_re", "coeff_im"],
    "formats": [np.int64, np.int32, np.int32, np.float64, np.float64],
    "offsets": [0, 8, 12, 16, 24],
    "itemsize": 32,
})


class QhipStream(C.Structure):
    _fields_ = [("ops", C.c_void_p), ("n_ops", C.c_int64), ("qubits", C.c_void_p),
                ("unitaries", C.c_void_p), ("dep_index", C.c_void_p), ("dep_alpha", C.c_void_p)]


class QhipObservable(C.Structure):
    _fields_ = [("terms", C.c_void_p), ("n_terms", C.c_int64), ("qubits", C.c_void_p),
                ("letters", C.c_void_p)]


class QhipConfig(C.Structure):
    _fields_ = [("fusion_max_qubits", C.c_int32), ("fusion_min_qubits", C.c_int32),
                ("tile_qubits", C.c_int32), ("low_qubits", C.c_int32),
                ("max_ops_per_pass", C.c_int32), ("use_mfma", C.c_int32),
                ("relabel_qubits", C.c_int32), ("stream_tiles", C.c_int32)]


class QhipApplyStats(C.Structure):
    _fields_ = [("fused_ops", C.c_int64), ("passes", C.c_int64), ("diagonal_ops", C.c_int64),
                ("mfma_ops", C.c_int64), ("flops", C.c_double), ("stream_passes", C.c_int64)]


class QhipBatchInfo(C.Structure):
    _fields_ = [("fwd_bytes", C.c_double), ("obs_bytes", C.c_double), ("rev_bytes", C.c_double),
                ("fwd_launches", C.c_int64), ("rev_launches", C.c_int64),
                ("instances_batched", C.c_int64), ("instances_single", C.c_int64)]


def _ptr(a: Optional[np.ndarray]) -> Optional[int]:
    return None if a is None or a.size == 0 else a.ctypes.data


@dataclass
class PackedStream:
    """A command stream as the C structs see it (buffers kept alive here)."""

    ops: np.ndarray
    qubits: np.ndarray
    unitaries: np.ndarray
    dep_index: np.ndarray
    dep_alpha: np.ndarray
    struct: QhipStream

    @property
    def ref(self):
        return C.byref(self.struct)


def pack_stream(commands: Sequence[Command], values: Optional[Mapping[str, float]] = None,
                param_index: Optional[Mapping[str, int]] = None) -> PackedStream:
    """Flatten commands into qhip_op records.

    ``values`` binds symbolic parameters; ``param_index`` (name -> gradient
    slot) fills the dep pools with d(angle)/d(parameter) for every tracked
    symbol, as gate_param_dependencies (qarp_simulator.cpp:1487-1507) does."""
    n = len(commands)
    ops = np.zeros(n, dtype=OP_DTYPE)
    # Fields are gathered into lists and written column by column (a
    # per-record structured write costs about a microsecond per field).
    gate, nq, qoff, npar, flags, cbit, uoff, doff, ndep = ([0] * n for _ in range(9))
    params = [(0.0, 0.0, 0.0, 0.0)] * n
    qubits: list[int] = []
    unitaries: list[np.ndarray] = []
    u_off = 0
    dep_idx: list[int] = []
    dep_alpha: list[float] = []
    custom = GateType.Custom
    slot = {s: int(j) for s, j in param_index.items()} if param_index is not None else None
    for i, cmd in enumerate(commands):
        g = cmd.gate
        qs = cmd.qubits
        gate[i] = int(g)   # IntEnum: int() is cheaper than the .value descriptor
        nq[i] = len(qs)
        qoff[i] = len(qubits)
        qubits.extend(qs)
        ps = cmd.params
        doff[i] = len(dep_idx)
        if ps:
            if len(ps) > 4:
                raise ValueError(f"gate '{g.name}' carries {len(ps)} parameters")
            vals = []
            nd = 0
            for p in ps:
                t = p._terms
                if p._fn is None:
                    if not t:
                        vals.append(p._offset)
                        continue
                    if values is None:
                        raise RuntimeError(f"gate '{g.name}' has unresolved symbolic parameter")
                    # Affine: Param.evaluate's arithmetic (the offset plus
                    # the sum of coefficient x value, summed from 0 in term
                    # order) in the same walk that reads the terms as the
                    # gradient coefficients.
                    acc = 0
                    try:
                        for s, c in t.items():
                            acc += c * float(values[s])
                            if slot is not None:
                                j = slot.get(s)
                                if j is not None:
                                    dep_idx.append(j)
                                    dep_alpha.append(float(c))
                                    nd += 1
                    except KeyError:
                        p.evaluate(values)   # raises naming the unbound symbols
                    vals.append(p._offset + acc)
                    continue
                # Compound: evaluated by its closure, coefficients probed.
                if values is None:
                    raise RuntimeError(f"gate '{g.name}' has unresolved symbolic parameter")
                vals.append(p.evaluate(values))
                if slot is not None:
                    try:
                        coeffs = affine_coefficients(p)
                    except CapabilityError as e:
                        raise CapabilityError(f"run_gradient: {e}") from None
                    for name, alpha in coeffs.items():
                        j = slot.get(name)
                        if j is not None:
                            dep_idx.append(j)
                            dep_alpha.append(float(alpha))
                            nd += 1
            npar[i] = len(vals)
            params[i] = tuple(vals) + (0.0,) * (4 - len(vals))
            ndep[i] = nd
        if cmd.condition_bits:
            flags[i] = QHIP_OP_CONDITIONAL
        cbit[i] = cmd.cbits[0] if cmd.cbits else -1
        if g == custom and cmd.unitary is not None:
            uoff[i] = u_off
            unitaries.append(np.ascontiguousarray(cmd.unitary, dtype=np.complex128).ravel())
            u_off += cmd.unitary.size
        else:
            uoff[i] = -1
    if n:
        ops["gate"] = gate
        ops["n_qubits"] = nq
        ops["qubit_offset"] = qoff
        ops["n_params"] = npar
        ops["params"] = np.asarray(params, dtype=np.float64)
        ops["flags"] = flags
        ops["cbit"] = cbit
        ops["unitary_offset"] = uoff
        ops["dep_offset"] = doff
        ops["n_deps"] = ndep
    q = np.asarray(qubits, dtype=np.uint32)
    u = np.concatenate(unitaries).view(np.float64) if unitaries else np.zeros(0, np.float64)
    di = np.asarray(dep_idx, dtype=np.int32)
    da = np.asarray(dep_alpha, dtype=np.float64)
    st = QhipStream(_ptr(ops), n, _ptr(q), _ptr(u), _ptr(di), _ptr(da))
    return PackedStream(ops, q, u, di, da, st)


Observable = Sequence[tuple]


@dataclass
class PackedObservable:
    terms: np.ndarray
    qubits: np.ndarray
    letters: np.ndarray
    struct: QhipObservable

    @property
    def ref(self):
        return C.byref(self.struct)


def pack_observable(observable: Observable) -> PackedObservable:
    """The observable ABI ``[([(qubit, 'X'|'Y'|'Z'|'I'), ...], coeff), ...]``
    (simulator/pauli_expectation.h:15-16)."""
    n = len(observable)
    terms = np.zeros(n, dtype=TERM_DTYPE)
    # Flattened with comprehensions and checked with numpy: a per-letter
    # Python loop costs about a millisecond per thousand strings.
    lens = [len(pauli) for pauli, _ in observable]
    qs = [q for pauli, _ in observable for q, _ in pauli]
    letters = [str(letter) for pauli, _ in observable for _, letter in pauli]
    q = np.asarray(qs, dtype=np.int64) if qs else np.zeros(0, np.int64)
    if q.size and q.min() < 0:
        raise ValueError(f"pauli_transition: observable qubit {int(q.min())} out of range")
    joined = "".join(letters)
    if len(joined) != len(letters):
        bad = next(s for s in letters if len(s) != 1)
        raise ValueError(f"pauli_transition: invalid Pauli '{bad}'")
    try:
        lt = np.frombuffer(joined.encode("ascii"), dtype=np.uint8).copy()
    except UnicodeEncodeError:
        bad = next(s for s in letters if not s.isascii())
        raise ValueError(f"pauli_transition: invalid Pauli '{bad}'") from None
    if n:   # whole columns at once (element-wise structured writes are slow)
        c = np.asarray([complex(coeff) for _, coeff in observable], dtype=np.complex128)
        offs = np.zeros(n, dtype=np.int64)
        if n > 1:
            np.cumsum(lens[:-1], out=offs[1:])
        terms["offset"] = offs
        terms["n_letters"] = lens
        terms["coeff_re"] = c.real
        terms["coeff_im"] = c.imag
    q32 = q.astype(np.uint32)
    st = QhipObservable(_ptr(terms), n, _ptr(q32), _ptr(lt))
    return PackedObservable(terms, q32, lt, st)


def config_struct(fusion_max_qubits: int, fusion_min_qubits: int, tile_qubits: int = 0,
                  low_qubits: int = 0, max_ops_per_pass: int = 0, use_mfma: int = -1,
                  relabel_qubits: int = -1, stream_tiles: int = -1) -> QhipConfig:
    return QhipConfig(fusion_max_qubits, fusion_min_qubits, tile_qubits, low_qubits,
                      max_ops_per_pass, use_mfma, relabel_qubits, stream_tiles)
