"""Structured programs: the typed kernels of ``qx.Program``
(src/qarpx/include/qarpx/simulator/program.h:17-83,
src/qarpx/src/simulator/program.cpp:486-605) that
``QarpSimulator.program_statevector`` / ``program_run`` execute on the device.

Kernels, in order: ``("gates", commands)`` through the fused apply,
``("permutation", qubits, table)`` through ``qhip_state_permute``,
``("dense", qubits, matrix)`` as one Custom op of the fused apply, and
``("controlled_powers", controls, exponents, targets, matrix)`` through
``qhip_state_controlled_powers``.  Kernels are validated on insertion with
the reference's messages, so execution never re-checks them.
"""
from __future__ import annotations

from typing import Mapping, Sequence

import numpy as np

from .ir import Command, GateType, as_commands, substitute_all

MAX_TABLE_QUBITS = 26        # program.cpp:35
MAX_DENSE_QUBITS = 8         # dense_kernel.h:13


def _check_distinct(qubits: Sequence[int], what: str) -> None:
    if len(set(qubits)) != len(qubits):
        raise ValueError(f"{what}: qubits must be distinct")


def _extract(x: np.ndarray, pos: Sequence[int]) -> np.ndarray:
    out = np.zeros_like(x)
    for b, p in enumerate(pos):
        out |= ((x >> np.uint64(p)) & np.uint64(1)) << np.uint64(b)
    return out


def _deposit(v: np.ndarray, pos: Sequence[int]) -> np.ndarray:
    out = np.zeros_like(v)
    for b, p in enumerate(pos):
        out |= ((v >> np.uint64(b)) & np.uint64(1)) << np.uint64(p)
    return out


def _compose(first: tuple, second: tuple):
    """The single permutation kernel equal to ``first`` then ``second`` over
    the sorted union of their qubits, or None past the table cap
    (compose, program.cpp:435-475)."""
    q1, t1 = first
    q2, t2 = second
    uni = sorted(set(q1) | set(q2))
    if len(uni) > MAX_TABLE_QUBITS:
        return None
    p1 = [uni.index(q) for q in q1]
    p2 = [uni.index(q) for q in q2]
    keep1 = np.uint64(~sum(1 << p for p in p1) & ((1 << 64) - 1))
    keep2 = np.uint64(~sum(1 << p for p in p2) & ((1 << 64) - 1))
    x = np.arange(1 << len(uni), dtype=np.uint64)
    x = (x & keep1) | _deposit(t1[_extract(x, p1)], p1)
    x = (x & keep2) | _deposit(t2[_extract(x, p2)], p2)
    return (uni, x)


class Program:
    """``qx.Program``: an execution-time lowering of one circuit into typed
    kernels."""

    def __init__(self):
        self._kernels: list = []

    def add_gates(self, commands: Sequence[Command]) -> None:
        self._kernels.append(("gates", as_commands(list(commands))))

    def add_permutation(self, qubits: Sequence[int], table) -> None:
        """|x> -> |table[x]> on qubits (local bit b <-> qubits[b]); table is a
        1-D integer bijection of range(2**len(qubits)).  Folds into a directly
        preceding permutation while their qubit union stays within the cap."""
        qs = [int(q) for q in qubits]
        _check_distinct(qs, "add_permutation")
        if not qs or len(qs) > MAX_TABLE_QUBITS:
            raise ValueError(f"add_permutation: 1..{MAX_TABLE_QUBITS} qubits required")
        t = np.asarray(table)
        n = 1 << len(qs)
        if t.ndim != 1 or t.size != n:
            raise ValueError(f"add_permutation: table length {t.size} for {len(qs)} qubit(s)")
        if np.any(t < 0) or np.any(t >= n):
            raise ValueError("add_permutation: table is not a bijection")
        t = t.astype(np.uint64)
        seen = np.zeros(n, dtype=bool)
        seen[t] = True
        if not seen.all():
            raise ValueError("add_permutation: table is not a bijection")
        nxt = (qs, t)
        if self._kernels and self._kernels[-1][0] == "permutation":
            merged = _compose(self._kernels[-1][1:], nxt)
            if merged is not None:
                self._kernels[-1] = ("permutation",) + merged
                return
        self._kernels.append(("permutation",) + nxt)

    def add_dense(self, qubits: Sequence[int], matrix) -> None:
        qs = [int(q) for q in qubits]
        _check_distinct(qs, "add_dense")
        if not qs or len(qs) > MAX_DENSE_QUBITS:
            raise ValueError(f"add_dense: 1..{MAX_DENSE_QUBITS} qubits required")
        m = np.asarray(matrix, dtype=np.complex128)
        d = 1 << len(qs)
        if m.shape != (d, d):
            raise ValueError("add_dense: matrix is not 2^k square")
        self._kernels.append(("dense", qs, m.copy()))

    def add_controlled_powers(self, controls: Sequence[int], exponents: Sequence[int],
                              targets: Sequence[int], matrix) -> None:
        """matrix**exponents[j] on targets wherever control j is |1>."""
        cs = [int(q) for q in controls]
        ts = [int(q) for q in targets]
        es = [int(e) for e in exponents]
        if not cs or not ts:
            raise ValueError("add_controlled_powers: controls and targets required")
        if len(es) != len(cs):
            raise ValueError("add_controlled_powers: one exponent per control")
        _check_distinct(cs + ts, "add_controlled_powers")
        m = np.asarray(matrix, dtype=np.complex128)
        d = 1 << len(ts)
        if len(ts) > MAX_TABLE_QUBITS or m.shape != (d, d):
            raise ValueError("add_controlled_powers: matrix is not 2^|targets| square")
        if any(e < 0 or e >= 1 << 64 for e in es):
            raise ValueError("add_controlled_powers: exponents must be unsigned 64-bit integers")
        self._kernels.append(("controlled_powers", cs, es, ts, m.copy()))

    def kernels(self) -> list:
        return list(self._kernels)

    def kinds(self) -> list[str]:
        return [k[0] for k in self._kernels]

    def __len__(self) -> int:
        return len(self._kernels)

    @property
    def min_register_width(self) -> int:
        top = 0
        for k in self._kernels:
            if k[0] == "gates":
                qs = [q for c in k[1] for q in c.qubits]
            elif k[0] == "controlled_powers":
                qs = list(k[1]) + list(k[3])
            else:
                qs = list(k[1])
            top = max([top] + [q + 1 for q in qs])
        return top

    def substituted(self, params: Mapping[str, float]) -> "Program":
        out = Program()
        out._kernels = [("gates", substitute_all(k[1], params)) if k[0] == "gates" else k
                        for k in self._kernels]
        return out

    def without_measurements(self) -> "Program":
        out = Program()
        out._kernels = [("gates", [c for c in k[1] if c.gate not in (GateType.Measure, GateType.Barrier)])
                        if k[0] == "gates" else k for k in self._kernels]
        return out


def needs_trajectory(c: Command) -> bool:
    """A command the sample-once path cannot take (qarp_simulator.cpp,
    any_needs_trajectory): Reset, a recorded Measure, a classical condition
    or a branch."""
    return (c.gate == GateType.Reset or (c.gate == GateType.Measure and bool(c.cbits))
            or bool(c.condition_bits) or c.gate == GateType.BranchBegin)


def measurements_terminal(commands: Sequence[Command]) -> bool:
    """measurements_are_terminal (qarp_simulator.cpp:226-252): no gate acts on
    a measured qubit after its measurement, and no reset, condition or
    branch appears."""
    measured = set()
    for c in commands:
        if c.gate == GateType.Reset or c.condition_bits:
            return False
        if c.gate in (GateType.BranchBegin, GateType.BranchElse, GateType.BranchEnd):
            return False
        if c.gate == GateType.Barrier:
            continue
        if c.gate == GateType.Measure:
            if c.cbits:
                measured.add(c.qubits[0])
            continue
        if any(q in measured for q in c.qubits):
            return False
    return True
