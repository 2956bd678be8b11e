"""Derivation helpers of the structured-program planner.

The reference exposes them as module functions of ``qarpx``
(src/qarpx/python/bindings.cpp:1302-1336, src/qarpx/src/simulator/program.cpp:
71-829) for its planner (src/qarp/_program.py:380-420), under the same
names here:

* ``_permutation_table(commands, qubits, max_rest, max_work_log2)`` — the
  basis permutation ``commands`` apply to ``qubits``, exact including phase,
  as an int64 table, or None (qhip_permutation_table, host computation);
* ``_local_unitary(commands, qubits)`` — the 2^k x 2^k unitary of
  ``commands`` on ``qubits`` (k <= 12), computed on the device: every basis
  state of the k qubits is a block of one 2k-qubit register
  (qhip_state_set_basis), the commands run once over all of them through the
  fused apply, and block c holds column c;
* ``_commands_digest``, ``_local_commands_digest``, ``_parts_digest`` — the
  64-bit FNV-1a cache keys of command streams (qhip_fnv1a_words mixes the
  words laid out here in the reference's order).

Commands may be this package's ``Command`` or the reference's ``qx.Command``
(read through gate, qubits, params, cbits, condition_bits,
condition_values, unitary).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

from . import _lib
from ._abi import pack_stream
from .ir import Command, GateType, as_commands

MAX_TABLE_QUBITS = 26      # program.cpp:35
MAX_UNITARY_QUBITS = 12    # program.cpp:38
_DIGEST_SEED = 0xCBF29CE484222325
_NON_UNITARY = {GateType.Measure, GateType.Reset, GateType.BranchBegin, GateType.BranchElse, GateType.BranchEnd}


def _check_distinct(qubits: Sequence[int], what: str) -> None:
    if len(set(qubits)) != len(qubits):
        raise ValueError(f"{what}: qubits must be distinct")


def _to_local(commands: Sequence[Command], qubits: Sequence[int]) -> Optional[list]:
    """``commands`` remapped onto range(len(qubits)), or None if any command
    is parametric, conditioned, non-unitary, or touches a qubit outside
    ``qubits`` (to_local, program.cpp:404-436)."""
    pos = {int(q): b for b, q in enumerate(qubits)}
    out = []
    for c in commands:
        if c.is_parametric() or c.condition_bits or c.cbits or c.gate in _NON_UNITARY:
            return None
        if any(q not in pos for q in c.qubits):
            return None
        out.append(Command(c.gate, qubits=[pos[q] for q in c.qubits], params=c.params, unitary=c.unitary))
    return out


def _permutation_table(commands, qubits: Sequence[int], max_rest: int, max_work_log2: int) -> Optional[np.ndarray]:
    """The basis permutation ``commands`` apply to ``qubits`` (local bit b
    <-> qubits[b]), exact including phase, as a 1-D int64 table, or None
    (permutation_table, program.cpp:714-724)."""
    qs = [int(q) for q in qubits]
    _check_distinct(qs, "permutation_table")
    if not qs or len(qs) > MAX_TABLE_QUBITS:
        return None
    local = _to_local(as_commands(commands), qs)
    if local is None:
        return None
    k = len(qs)
    ps = pack_stream(local)
    found = C.c_int32(0)
    table = np.zeros(1 << k, dtype=np.uint64)
    _lib.check(_lib.load().qhip_permutation_table(ps.ref, k, int(max_rest), int(max_work_log2), C.byref(found),
                                                  table.ctypes.data))
    return table.astype(np.int64) if found.value else None


def _local_unitary(commands, qubits: Sequence[int], simulator=None) -> np.ndarray:
    """The 2^k x 2^k unitary ``commands`` apply to ``qubits`` (local bit b <->
    qubits[b]; local_unitary_of, program.cpp:726-738), k <= 12."""
    from .simulator import DeviceState, QarpSimulator
    qs = [int(q) for q in qubits]
    _check_distinct(qs, "local_unitary_of")
    if not qs or len(qs) > MAX_UNITARY_QUBITS:
        raise ValueError(f"local_unitary_of: 1..{MAX_UNITARY_QUBITS} qubits required")
    local = _to_local(as_commands(commands), qs)
    if local is None:
        raise ValueError("local_unitary_of: commands must be concrete, unconditioned and unitary, "
                         "on the listed qubits only")
    k = len(qs)
    d = 1 << k
    sim = simulator if simulator is not None else QarpSimulator()
    st = DeviceState(2 * k, "complex128")
    labels = np.arange(d, dtype=np.uint64)
    _lib.check(_lib.load().qhip_state_set_basis(st.handle, labels.ctypes.data, d, d))
    if local:
        sim.apply(local, st)
    # Amplitude r + d c is U[r, c]: block c is column c.
    return np.asfortranarray(st.download().reshape(d, d).T)


# ── digests ──────────────────────────────────────────────────────────────────

def _f64_bits(x: float) -> int:
    return int(np.float64(x).view(np.uint64))


def _command_words(c, qmap: Optional[dict]) -> list:
    """One command as the 64-bit words the reference mixes (mix_command,
    program.cpp:744-772): gate, qubits (through qmap), cbits, conditions,
    parameters (a concrete value's bits, else the characters of its string
    form: repr() of the caller's parameter object, the reference's
    to_string for its own Param), and a Custom unitary's entries."""
    gate = int(getattr(c.gate, "value", c.gate))
    qubits = [int(q) for q in c.qubits]
    words = [gate, len(qubits)] + [qmap[q] if qmap is not None else q for q in qubits]
    cbits = [int(b) for b in getattr(c, "cbits", ())]
    words += [len(cbits)] + cbits
    cb = [int(b) for b in getattr(c, "condition_bits", ())]
    cv = [bool(v) for v in getattr(c, "condition_values", ())]
    words.append(len(cb))
    for b, v in zip(cb, cv):
        words += [b, 1 if v else 0]
    params = list(getattr(c, "params", ()))
    words.append(len(params))
    for p in params:
        if isinstance(p, (int, float, np.integer, np.floating)):
            words.append(_f64_bits(float(p)))
        elif p.is_concrete():
            words.append(_f64_bits(float(p.value())))
        else:
            words += [b for b in repr(p).encode()]
    u = getattr(c, "unitary", None)
    if u is not None:
        m = np.asarray(u, dtype=np.complex128)
        pairs = np.stack([m.real, m.imag], axis=-1).reshape(-1)
        words += pairs.view(np.uint64).tolist()
    return words


def _mix(words: list, seed: int = _DIGEST_SEED) -> int:
    arr = np.ascontiguousarray(np.array(words, dtype=np.uint64))
    out = C.c_uint64(0)
    _lib.check(_lib.load().qhip_fnv1a_words(arr.ctypes.data if arr.size else None, arr.size, seed, C.byref(out)))
    return int(out.value)


def _commands_digest(commands) -> int:
    """64-bit digest of a command stream (commands_digest, program.cpp:799-803)."""
    words = []
    for c in commands:
        words += _command_words(c, None)
    return _mix(words)


def _local_commands_digest(commands, qubits: Sequence[int]) -> int:
    """The digest of ``commands`` remapped onto range(len(qubits)); the plain
    digest when a command lies outside ``qubits`` (local_commands_digest,
    program.cpp:782-797)."""
    qmap = {int(q): b for b, q in enumerate(qubits)}
    if any(int(q) not in qmap for c in commands for q in c.qubits):
        return _commands_digest(commands)
    words = []
    for c in commands:
        words += _command_words(c, qmap)
    return _mix(words)


def _parts_digest(parts) -> int:
    """The digest of the stream in which each part's commands repeat
    ``count`` times, in order (parts_digest, program.cpp:805-811)."""
    words = []
    for commands, count in parts:
        one = []
        for c in commands:
            one += _command_words(c, None)
        words += one * int(count)
    return _mix(words)


permutation_table = _permutation_table
local_unitary = _local_unitary
commands_digest = _commands_digest
local_commands_digest = _local_commands_digest
parts_digest = _parts_digest
