"""Transition columns U† H U |s> for batches of basis states s: the
``TransitionColumnProvider`` of the reference
(src/qarpx/include/qarpx/simulator/transition_columns.h:43-63,
src/qarpx/src/simulator/transition_columns.cpp:22-48, .cu:236-380).

A batch of B = 2^b sources lives in one (n + b)-qubit device state, source
j in the block of amplitudes [j·2^n, (j+1)·2^n): U and U† are compiled once
as plans on that register (their gates touch qubits < n, so every pass
advances all B columns together), and H acts through the Pauli-sum kernel
on the same register.  One batch is: basis states (qhip_state_set_basis),
U, H into a second state, U†, and one copy of the B columns to the host.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from typing import Optional, Sequence

import numpy as np

from . import _lib
from ._abi import pack_observable
from .errors import CapabilityError
from .ir import Command, GateType

_NON_UNITARY = (GateType.Reset, GateType.Measure, GateType.BranchBegin, GateType.BranchElse, GateType.BranchEnd)


def _validate(U: Sequence[Command], n: int, observable) -> list:
    """CudaqSimulator::make_transition_column_provider's checks
    (transition_columns.cpp:28-46) and those of the provider's term upload
    (device_terms, transition_columns.cu:141-176): concrete finite
    parameters, unconditional measurement-free gates on the register, finite
    real coefficients, X/Y/Z letters on distinct qubits of the register."""
    for c in U:
        for p in c.params:
            if not p.is_concrete() or not math.isfinite(p.value()):
                raise ValueError("Column commands require finite concrete parameters")
        if c.gate in _NON_UNITARY or c.condition_bits or c.cbits:
            raise CapabilityError("Columns require unconditional measurement-free unitaries")
        if any(q >= n for q in c.qubits):
            raise ValueError("TransitionColumnProvider: U acts outside the register")
    terms = []
    for letters, coeff in observable:
        z = complex(coeff)
        if not (math.isfinite(z.real) and math.isfinite(z.imag)) or z.imag != 0:
            raise ValueError("Column observable coefficients must be finite and real")
        used = set()
        for q, op in letters:
            if int(q) >= n:
                raise ValueError("TransitionColumnProvider: Hamiltonian acts outside the register")
            if int(q) in used:
                raise ValueError("TransitionColumnProvider Pauli term repeats a qubit")
            used.add(int(q))
            if op not in ("X", "Y", "Z"):
                raise ValueError("TransitionColumnProvider: Pauli terms must use X, Y, or Z")
        terms.append((list(letters), z.real))
    return terms


class TransitionColumnProvider:
    """Columns of U† H U at basis-state sources, row layout [source, target]."""

    def __init__(self, sim, U: Sequence[Command], n: int, observable, initial_batch_capacity: int = 1,
                 sector_target_labels: Optional[np.ndarray] = None):
        from .simulator import CompiledCircuit, DeviceState
        n = int(n)
        terms = _validate(U, n, observable)
        if n < 1 or n >= 63:
            raise ValueError("TransitionColumnProvider requires 1 <= n_qubits < 63")
        if int(initial_batch_capacity) <= 0:
            raise ValueError("TransitionColumnProvider requires positive batch capacity")
        self._n = n
        self._dim = 1 << n
        sector = np.zeros(0, np.uint64) if sector_target_labels is None else np.asarray(sector_target_labels)
        sector = sector.astype(np.uint64)
        if sector.size and int(sector.max()) >= self._dim:
            raise ValueError("TransitionColumnProvider sector target is outside the register")
        if sector.size > 1 and np.any(sector[1:] <= sector[:-1]):
            raise ValueError("TransitionColumnProvider sector targets must be strictly increasing")
        self._sector = sector
        self._b = max(0, (int(initial_batch_capacity) - 1).bit_length())
        self._batch = 1 << self._b
        width = n + self._b
        if width > 40:
            raise CapabilityError(f"TransitionColumnProvider: a batch of {self._batch} columns of {n} qubits "
                                  "exceeds a 40-qubit device register")
        t0 = time.perf_counter()
        self._sim = sim
        self._obs = pack_observable([(t, c) for t, c in terms])
        self._has_terms = bool(terms)
        # A global phase cancels between U and its adjoint (gate_matrices drops it).
        gates = [c for c in U if c.gate not in (GateType.GPhase, GateType.Barrier)]
        dagger = [c.dagger() for c in reversed(gates)]
        self._u = CompiledCircuit(sim, gates, width)
        self._ud = CompiledCircuit(sim, dagger, width)
        self._psi = DeviceState(width, sim.precision)
        self._phi = DeviceState(width, sim.precision)
        self._diag = {k: 0 for k in (
            "column_batches", "columns", "capacity_growths", "handle_creations", "term_uploads", "gate_uploads",
            "transferred_bytes", "compact_entries", "pruned_entries", "sector_omitted_entries")}
        self._diag.update({"sector_target_count": int(sector.size), "capacity": self._batch,
                           "state_buffer_bytes": 2 * (1 << width) * (16 if sim.precision == "complex128" else 8),
                           "workspace_bytes": 0, "term_uploads": 1, "gate_uploads": 2})
        for k in ("basis_wall_seconds", "u_wall_seconds", "hamiltonian_wall_seconds", "u_dagger_wall_seconds",
                  "transfer_wall_seconds", "sector_gather_wall_seconds", "compaction_wall_seconds",
                  "max_pruned_amplitude"):
            self._diag[k] = 0.0
        self._diag["allocation_wall_seconds"] = time.perf_counter() - t0

    def n_qubits(self) -> int:
        return self._n

    @property
    def diagnostics(self) -> dict:
        return dict(self._diag)

    def _check_sources(self, source_labels) -> np.ndarray:
        src = np.asarray(source_labels).astype(np.uint64).ravel()
        if src.size and int(src.max()) >= self._dim:
            raise ValueError("TransitionColumnProvider source label is outside the register")
        if self._sector.size and src.size and not np.all(np.isin(src, self._sector)):
            raise ValueError("TransitionColumnProvider source label is outside the configured sector")
        return src

    def _evaluate(self, part: np.ndarray) -> None:
        """U† H U on the basis states `part` (at most one batch): the result
        is left in the second state, block j = column of part[j]."""
        lib = _lib.load()
        d = self._diag
        t = time.perf_counter()
        _lib.check(lib.qhip_state_set_basis(self._psi.handle, part.ctypes.data, part.size, self._dim))
        t1 = time.perf_counter()
        self._u.apply(self._psi)
        _lib.check(lib.qhip_synchronize())
        t2 = time.perf_counter()
        if self._has_terms:
            _lib.check(lib.qhip_apply_observable(self._psi.handle, self._obs.ref, self._phi.handle))
        else:
            _lib.check(lib.qhip_state_set_basis(self._phi.handle, None, 0, self._dim))
        _lib.check(lib.qhip_synchronize())
        t3 = time.perf_counter()
        self._ud.apply(self._phi)
        _lib.check(lib.qhip_synchronize())
        t4 = time.perf_counter()
        d["basis_wall_seconds"] += t1 - t
        d["u_wall_seconds"] += t2 - t1
        d["hamiltonian_wall_seconds"] += t3 - t2
        d["u_dagger_wall_seconds"] += t4 - t3
        d["column_batches"] += 1

    def _rows(self, src: np.ndarray, gather: bool) -> np.ndarray:
        width = len(self._sector) if gather else self._dim
        out = np.empty((src.size, width), dtype=np.complex128)
        d = self._diag
        for start in range(0, src.size, self._batch):
            part = np.ascontiguousarray(src[start:start + self._batch])
            self._evaluate(part)
            t4 = time.perf_counter()
            if not gather and part.size == self._batch:
                # A full batch lands in its output rows directly.
                self._phi.download(out[start:start + part.size].reshape(-1))
                t5 = time.perf_counter()
            else:
                rows = self._phi.download().reshape(self._batch, self._dim)[:part.size]
                t5 = time.perf_counter()
                out[start:start + part.size] = rows[:, self._sector.astype(np.int64)] if gather else rows
            t6 = time.perf_counter()
            d["transfer_wall_seconds"] += t5 - t4
            d["sector_gather_wall_seconds"] += t6 - t5 if gather else 0.0
            d["transferred_bytes"] += self._batch * self._dim * 16
        d["columns"] += int(src.size)
        return out

    def columns(self, source_labels) -> np.ndarray:
        """(len(sources), 2^n) complex128: row j is U† H U |sources[j]>."""
        src = self._check_sources(source_labels)
        if src.size == 0:
            return np.zeros((0, self._dim), dtype=np.complex128)
        return self._rows(src, False)

    def _compact_on_device(self, src: np.ndarray, tol: float):
        """Full columns compacted on the device (qhip_compact): only the kept
        entries cross to the host.  The output arrays hold the entries kept so
        far plus room for one whole batch: after the first batch they are
        sized from its kept fraction for all remaining batches, and grow by
        half when a later batch needs more, so host memory follows the kept
        entries, not the dense columns.  Per batch: one qhip_compact call
        (which also returns the largest pruned magnitude), row offsets from
        the ascending keys by binary search, labels masked in place.
        Non-finite entries: the batch's FP64 norm is finite whenever every
        amplitude is (short of overflow); otherwise the batch is downloaded
        and checked entry by entry, pruned entries included, as the
        reference's loop does (transition_columns.cu:266-276)."""
        lib = _lib.load()
        d = self._diag
        dense = src.size * self._dim
        cap = min(dense, self._batch * self._dim)
        keys = np.empty(cap, dtype=np.int64)
        vals = np.empty(cap, dtype=np.complex128)
        offsets = np.zeros(src.size + 1, dtype=np.uint64)
        pos = 0
        max_pruned = C.c_double()
        for start in range(0, src.size, self._batch):
            part = np.ascontiguousarray(src[start:start + self._batch])
            n_amps = part.size * self._dim
            if pos + n_amps > cap:
                # Sized from the kept fraction so far (one resize for columns
                # of even density), never below one more batch nor above dense.
                est = int(pos / (start * self._dim) * dense * 1.05) + n_amps if start else 0
                new_cap = min(dense, max(pos + n_amps, est, cap + cap // 2))
                keys = np.concatenate([keys[:pos], np.empty(new_cap - pos, np.int64)])
                vals = np.concatenate([vals[:pos], np.empty(new_cap - pos, np.complex128)])
                cap = new_cap
            self._evaluate(part)
            t = time.perf_counter()
            if not math.isfinite(self._phi.norm2()):
                rows = self._phi.download()[:n_amps]
                if not np.isfinite(np.hypot(rows.real, rows.imag)).all():
                    raise RuntimeError("Exact column contains a nonfinite coefficient")
            cnt = C.c_int64()
            _lib.check(lib.qhip_compact(self._phi.handle, n_amps, tol, keys[pos:].ctypes.data,
                                        vals[pos:].ctypes.data, cap - pos, C.byref(cnt), C.byref(max_pruned)))
            d["max_pruned_amplitude"] = max(d["max_pruned_amplitude"], max_pruned.value)
            k = keys[pos:pos + cnt.value]
            ends = np.arange(1, part.size + 1, dtype=np.int64) << self._n
            offsets[start + 1:start + part.size + 1] = pos + np.searchsorted(k, ends)
            np.bitwise_and(k, self._dim - 1, out=k)
            pos += cnt.value
            d["transferred_bytes"] += 24 * cnt.value
            d["compaction_wall_seconds"] += time.perf_counter() - t
        d["columns"] += int(src.size)
        d["compact_entries"] += pos
        d["pruned_entries"] += int(src.size) * self._dim - pos
        return offsets, keys[:pos].view(np.uint64), vals[:pos]

    def compact_columns(self, source_labels, absolute_tolerance: float):
        """(offsets, labels, values): per source the entries of magnitude above
        the tolerance (restricted to the sector targets when one is set), in
        target order; offsets has len(sources) + 1 entries."""
        tol = float(absolute_tolerance)
        if not math.isfinite(tol) or tol < 0.0:
            raise ValueError("TransitionColumnProvider requires a finite, non-negative compaction tolerance")
        src = self._check_sources(source_labels)
        gather = self._sector.size > 0
        if src.size == 0:
            return np.zeros(1, np.uint64), np.zeros(0, np.uint64), np.zeros(0, np.complex128)
        if not gather:
            return self._compact_on_device(src, tol)
        rows = self._rows(src, gather)
        t = time.perf_counter()
        targets = self._sector if gather else np.arange(self._dim, dtype=np.uint64)
        mag = np.abs(rows)
        if not np.isfinite(mag).all():
            raise RuntimeError("Exact column contains a nonfinite coefficient")
        keep = mag > tol
        r, c = np.nonzero(keep)          # row-major: per source, targets ascending
        offsets = np.zeros(src.size + 1, dtype=np.uint64)
        offsets[1:] = np.cumsum(np.bincount(r, minlength=src.size))
        labels = targets[c].astype(np.uint64)
        values = rows[r, c].astype(np.complex128)
        pruned = mag[~keep]
        d = self._diag
        d["max_pruned_amplitude"] = max(d["max_pruned_amplitude"], float(pruned.max()) if pruned.size else 0.0)
        d["compact_entries"] += int(labels.size)
        d["pruned_entries"] += int(src.size) * rows.shape[1] - int(labels.size)
        if gather:
            d["sector_omitted_entries"] += int(src.size) * (self._dim - rows.shape[1])
        d["compaction_wall_seconds"] += time.perf_counter() - t
        return offsets, labels, values


def make_transition_column_provider(sim, U: Sequence[Command], n: int, observable,
                                    initial_batch_capacity: int = 1,
                                    sector_target_labels: Optional[np.ndarray] = None) -> TransitionColumnProvider:
    """CudaqSimulator.make_transition_column_provider (bindings_montecarlo.cpp:375-381)."""
    return TransitionColumnProvider(sim, U, n, observable, initial_batch_capacity, sector_target_labels)
