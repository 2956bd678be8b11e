"""Primitive counterparts of the reference's StateVector, Sampler and
PauliAveraging, run on the device through :class:`QarpSimulator`.

Only what the statevector hot path feeds is kept: the circuits are plain
command streams (the reference's block trees are out of scope) and the
observable is the reference's ABI, ``[([(qubit, 'X'|'Y'|'Z'|'I'), ...],
coefficient), ...]``.

* :class:`StateVector` — ``run_from_amplitudes``
  (src/qarp/algorithms/_primitives/state_vector.py:361-431): EXPECTATION_VALUE
  as ⟨ψ|H|ψ⟩, OVERLAP as ⟨bra|ket⟩, TRANSITION_AMPLITUDE as ⟨bra|H|ket⟩.  The
  states stay on the device; a Z-only operator is one x-group of the device
  transition kernel (a single pass over |ψ|²), which is what the reference's
  Walsh-Hadamard diagonal computes on the host (:156-201).
* :class:`Sampler` — terminal measurement of all or some qubits: COUNTS via
  the device CDF and search (bit-exact with the reference for the same seed),
  EXACT via the device |ψ|² compaction with the 1e-12 prune
  (src/qarp/engines/_engine.py:79-116); results are marginalised by
  :func:`distribution_from_result`.
* :class:`PauliAveraging` — ``circuits`` / ``run``
  (src/qarp/algorithms/_primitives/pauli_averaging.py:119-205): one
  measure-all circuit per group of terms behind a basis change, and the
  parity estimator E[term] = s·Σ_outcome w·(-1)^popcount(outcome & mask).
  Groups here are qubit-wise commuting (first fit in term order) with the
  basis change H for X and S†·H for Y; the reference's default groups by
  general commutation with a synthesised Clifford.  Both estimate the same
  ⟨H⟩ (identically so for EXACT).
"""
from __future__ import annotations

import enum
from typing import Mapping, Optional, Sequence, Union

import numpy as np

from .errors import CapabilityError
from .grouping import FullyCommuting, GroupingStrategy, diagonalise_group
from .ir import Command, GateType, as_commands
from .results import (EXACT, EXACT_PRUNE_TOL, ExactResult, SamplingDistribution, Shots,
                      distribution_from_result, outcome_arrays)
from .simulator import QarpSimulator

Observable = Sequence[tuple]


class Target(enum.Enum):
    EXPECTATION_VALUE = "expectation_value"
    OVERLAP = "overlap"
    TRANSITION_AMPLITUDE = "transition_amplitude"


def free_symbols(commands: Sequence[Command]) -> list:
    out: dict = {}
    for c in commands:
        for p in c.params:
            for s in p.free_symbols():
                out.setdefault(s, None)
    return list(out)


def bind(commands: Sequence[Command], params: Optional[Mapping[str, float]]) -> list:
    if not params:
        return list(commands)
    return [c.substitute(params) if c.is_parametric() else c for c in commands]


def is_z_only(observable: Optional[Observable]) -> bool:
    return observable is not None and all(p in ("Z", "I") for term, _ in observable for _, p in term)


def diagonal_pauli_vector(observable: Observable, n_qubits: int) -> np.ndarray:
    """Host diagonal of a Z-only operator: the Walsh-Hadamard transform of the
    coefficients placed at their Z-masks (state_vector.py:156-201).  Used to
    check the device path; O(n 2^n)."""
    if not is_z_only(observable):
        raise ValueError("diagonal_pauli_vector: operator has X or Y terms")
    dim = 1 << n_qubits
    vals = np.zeros(dim, dtype=np.complex128)
    for term, coeff in observable:
        mask = 0
        for q, p in term:
            if q >= n_qubits:
                raise ValueError(f"diagonal_pauli_vector: term acts on qubit {q} but n_qubits is {n_qubits}")
            if p == "Z":
                mask |= 1 << q
        vals[mask] += complex(coeff)
    half = 1
    while half < dim:
        v = vals.reshape(-1, 2, half)
        lo = v[:, 0, :].copy()
        v[:, 0, :] += v[:, 1, :]
        v[:, 1, :] = lo - v[:, 1, :]
        half *= 2
    return vals.real.copy() if not vals.imag.any() else vals


class StateVector:
    """Amplitude primitive; ``run`` returns float for EXPECTATION_VALUE (the
    operator is Hermitian) and complex for OVERLAP / TRANSITION_AMPLITUDE."""

    gradient_kind = "expectation"

    def __init__(self, ket: Sequence[Command], n_qubits: int, observable: Optional[Observable] = None, *,
                 bra: Optional[Sequence[Command]] = None,
                 target: Target = Target.EXPECTATION_VALUE, initial_state=None):
        self.ket = as_commands(list(ket))
        self.bra = as_commands(list(bra)) if bra is not None else None
        self.n_qubits = int(n_qubits)
        self.observable = observable
        self.target = Target(target)
        self.initial_state = initial_state
        if self.target is Target.EXPECTATION_VALUE and observable is None:
            raise ValueError("StateVector: EXPECTATION_VALUE needs an observable")
        if self.target is Target.TRANSITION_AMPLITUDE and (observable is None or bra is None):
            raise ValueError("StateVector: TRANSITION_AMPLITUDE needs an observable and a bra circuit")
        if self.target is Target.OVERLAP and bra is None:
            raise ValueError("StateVector: OVERLAP needs a bra circuit")
        if self.target is not Target.EXPECTATION_VALUE:
            self.gradient_kind = "none"

    def symbols(self) -> list:
        return free_symbols(self.ket + (self.bra or []))

    def _ket_state(self, sim: QarpSimulator, params):
        st = sim.new_state(self.n_qubits, self.initial_state, "statevector")
        sim.apply(bind(self.ket, params), st)
        return st

    def _bra_state(self, sim: QarpSimulator, params):
        st = sim.new_state(self.n_qubits)
        sim.apply(bind(self.bra, params), st)
        return st

    def run(self, sim: QarpSimulator, params: Optional[Mapping[str, float]] = None) -> Union[float, complex]:
        ket = self._ket_state(sim, params)
        if self.target is Target.EXPECTATION_VALUE:
            return sim.expectation_state(ket, self.observable)
        bra = self._bra_state(sim, params)
        if self.target is Target.OVERLAP:
            return sim.transition_state(bra, ket, [([], 1.0)])
        return sim.transition_state(bra, ket, self.observable)


def _measured_stream(commands: Sequence[Command], measured: Sequence[int]) -> list:
    # Unrecorded terminal measures: the counts stay full-register outcomes,
    # marginalised afterwards.
    body = [c for c in commands if c.gate not in (GateType.Measure, GateType.Barrier)]
    return body + [Command(GateType.Measure, q) for q in measured]


class Sampler:
    """Terminal-measurement sampling of ``measured`` (default: every qubit)."""

    gradient_kind = "none"

    def __init__(self, circuit: Sequence[Command], n_qubits: int, n_shots: Union[int, Shots] = 1024,
                 measured: Optional[Sequence[int]] = None, initial_state=None):
        self.circuit = as_commands(list(circuit))
        self.n_qubits = int(n_qubits)
        if not (n_shots is EXACT or (isinstance(n_shots, (int, np.integer)) and n_shots >= 0)):
            raise ValueError(f"n_shots must be a non-negative int or EXACT, got {n_shots!r}")
        self.n_shots = n_shots
        self.measured = list(range(self.n_qubits)) if measured is None else [int(q) for q in measured]
        self.initial_state = initial_state

    def symbols(self) -> list:
        return free_symbols(self.circuit)

    def raw(self, sim: QarpSimulator, params=None, seed: Optional[int] = None):
        """The full-register result: ExactResult (EXACT) or SamplingResult."""
        if self.n_shots is EXACT:
            st = sim.new_state(self.n_qubits, self.initial_state, "run")
            body = [c for c in bind(self.circuit, params) if c.gate not in (GateType.Measure, GateType.Barrier)]
            sim.apply(body, st)
            keys, probs = sim.exact_state(st, EXACT_PRUNE_TOL)
            return ExactResult(self.n_qubits, keys, probs)
        stream = _measured_stream(bind(self.circuit, params), self.measured)
        return sim.run(stream, self.n_qubits, int(self.n_shots), seed, initial_state=self.initial_state)

    def run(self, sim: QarpSimulator, params=None, seed: Optional[int] = None) -> SamplingDistribution:
        return distribution_from_result(self.raw(sim, params, seed), self.measured)


def decompose_observable(observable: Observable) -> tuple:
    """(constant, terms, coeffs) of a Hermitian observable: identity terms
    summed into the constant, the others as {qubit: letter} dicts with their
    real coefficients (PauliAveraging._decompose_operator,
    pauli_averaging.py:100-118, which keeps the real part)."""
    constant = 0.0
    terms, coeffs = [], []
    for term, coeff in observable:
        letters: dict = {}
        seen = set()
        for q, p in term:
            q, p = int(q), str(p)
            if len(p) != 1 or p not in "IXYZ":
                raise ValueError(f"PauliAveraging: invalid Pauli '{p}'")
            if q in seen:
                raise ValueError(f"PauliAveraging: qubit {q} repeated within one Pauli string")
            seen.add(q)
            if p != "I":
                letters[q] = p
        c = complex(coeff).real
        if not letters:
            constant += c
            continue
        terms.append(letters)
        coeffs.append(c)
    return constant, terms, np.asarray(coeffs, dtype=np.float64)


def _parity(x: np.ndarray) -> np.ndarray:
    x = x ^ (x >> 32)
    x = x ^ (x >> 16)
    x = x ^ (x >> 8)
    x = x ^ (x >> 4)
    x = x ^ (x >> 2)
    x = x ^ (x >> 1)
    return (x & 1).astype(np.float64)


class PauliAveraging:
    """⟨ψ|H|ψ⟩ estimated from the measurement counts of one circuit per group
    of commuting terms: ket → the group's diagonalising Clifford → measure
    every qubit (pauli_averaging.py:119-205).  ``grouping`` defaults to
    FullyCommuting, as the reference's does."""

    gradient_kind = "expectation"

    def __init__(self, ket: Sequence[Command], n_qubits: int, observable: Observable,
                 n_shots: Union[int, Shots] = 1024, grouping: Optional[GroupingStrategy] = None):
        self.ket = as_commands(list(ket))
        self.n_qubits = int(n_qubits)
        self.observable = list(observable)
        self.n_shots = n_shots
        self.grouping = grouping if grouping is not None else FullyCommuting()
        self.constant, self._terms, self._coeffs = decompose_observable(self.observable)
        for t in self._terms:
            if max(t) >= self.n_qubits:
                raise ValueError(f"PauliAveraging: observable qubit {max(t)} out of range for "
                                 f"{self.n_qubits} qubits")
        self.groups = []   # (term indices, Clifford commands, z masks, signs)
        for grp in (self.grouping.group(self._terms, self.n_qubits) if self._terms else []):
            clifford, masks, signs = diagonalise_group([self._terms[i] for i in grp], self.n_qubits)
            self.groups.append((list(grp), clifford, masks, signs))

    @property
    def n_groups(self) -> int:
        return len(self.groups)

    @property
    def n_terms(self) -> int:
        return len(self._terms)

    def symbols(self) -> list:
        return free_symbols(self.ket)

    def group_circuit(self, g: int) -> list:
        """The reference's circuit of group g without its measure layer:
        ket, then the group's Clifford."""
        return self.ket + list(self.groups[g][1])

    def run(self, sim: QarpSimulator, params=None, seeds: Optional[Sequence[Optional[int]]] = None) -> float:
        """The ket runs once into a device state; each group copies it,
        applies only its Clifford and samples the copy (the reference runs
        ket + Clifford per group: the same state up to rounding, with the same
        per-group seeds)."""
        energy = self.constant
        if not self.groups:
            return float(energy)
        ket = bind(self.ket, params)
        shared = not any(c.gate == GateType.Reset or c.condition_bits or (c.gate == GateType.Measure and c.cbits)
                         or c.gate in (GateType.BranchBegin, GateType.BranchElse, GateType.BranchEnd)
                         for c in ket)
        if shared:
            base = sim.new_state(self.n_qubits)
            sim.apply([c for c in ket if c.gate not in (GateType.Measure, GateType.Barrier)], base)
            work = sim.new_state(self.n_qubits)
        for g, (idx, clifford, masks, signs) in enumerate(self.groups):
            seed = None if seeds is None else seeds[g]
            if not shared:   # trajectory commands: the per-group circuit raises as run() does
                res = Sampler(self.group_circuit(g), self.n_qubits, self.n_shots).raw(sim, params, seed)
            else:
                work.copy_from(base)
                if clifford:
                    sim.apply(clifford, work)
                if self.n_shots is EXACT:
                    keys, probs = sim.exact_state(work, EXACT_PRUNE_TOL)
                    res = ExactResult(self.n_qubits, keys, probs)
                else:
                    res = sim._terminal_result(work, [], self.n_qubits, int(self.n_shots), seed)
            energy += self.group_estimate(res, g)
        return float(energy)

    def group_estimate(self, res, g: int) -> float:
        """Σ_t c_t ⟨P_t⟩ of group g from its counts (or exact distribution):
        after the Clifford, term t is s_t Z^{mask_t}, so its expectation is
        s_t Σ_outcome w (-1)^popcount(outcome & mask_t) / n_shots."""
        idx, _, masks, signs = self.groups[g]
        outcomes, weights = outcome_arrays(res)
        m = np.asarray(masks, dtype=np.int64)
        sg = np.where(signs, -1.0, 1.0)
        parity = _parity(outcomes[:, None] & m[None, :])
        e_terms = sg * (weights @ (1.0 - 2.0 * parity)) / res.n_shots
        return float(self._coeffs[idx] @ e_terms)

    def __repr__(self) -> str:
        return (f"PauliAveraging(n_terms={self.n_terms}, n_groups={self.n_groups}, n_shots={self.n_shots}, "
                f"grouping={self.grouping!r})")


def require_capability(prim, method: str) -> None:
    if getattr(prim, "gradient_kind", "none") == "none":
        raise CapabilityError(f"{type(prim).__name__} ({getattr(prim, 'target', '')}) has no {method} gradient")
