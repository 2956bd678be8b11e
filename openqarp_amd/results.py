"""Result types of the sampling primitives.

``EXACT`` selects the infinite-shot Born distribution (``Shots.EXACT``,
src/qarp/_types.py:30-32); :class:`ExactResult` carries it as sorted int64
outcome keys and probabilities with ``n_shots = 1`` so that every estimator's
``count / n_shots`` reads a probability (src/qarp/_types.py:35-58).
:func:`distribution_from_result` marginalises a counts-shaped result onto the
measured qubits, keys packed LSB-first (src/qarp/_sampling_distribution.py:
41-97).
"""
from __future__ import annotations

import enum
from collections.abc import Mapping
from dataclasses import dataclass, field
from functools import cached_property
from typing import Iterator, Optional, Sequence

import numpy as np


class Shots(enum.Enum):
    EXACT = "exact"


EXACT = Shots.EXACT

# Born probabilities at or below this are dropped from an EXACT result
# (src/qarp/engines/_engine.py:76).
EXACT_PRUNE_TOL = 1e-12


@dataclass
class ExactResult:
    n_qubits: int
    keys: np.ndarray      # int64 outcome bitmasks, ascending
    probs: np.ndarray     # float64
    n_shots: int = 1
    is_exact: bool = field(default=True, init=False)

    @cached_property
    def counts(self) -> dict:
        return dict(zip(self.keys.tolist(), self.probs.tolist()))


def outcome_arrays(result) -> tuple:
    """(outcomes int64, weights float64) of a counts-shaped result."""
    keys = getattr(result, "keys", None)
    if keys is not None and getattr(result, "is_exact", False):
        return np.asarray(keys, dtype=np.int64), np.asarray(result.probs, dtype=np.float64)
    counts = result.counts
    out = np.fromiter(counts.keys(), dtype=np.int64, count=len(counts))
    w = np.fromiter(counts.values(), dtype=np.float64, count=len(counts))
    return out, w


def pack_bits(outcomes: np.ndarray, positions: Sequence[int]) -> np.ndarray:
    """Bit ``positions[i]`` of every outcome moved to bit i; other bits dropped."""
    positions = list(positions)
    outcomes = np.asarray(outcomes, dtype=np.int64)
    if positions == list(range(len(positions))):
        return outcomes & ((1 << len(positions)) - 1)
    packed = np.zeros_like(outcomes)
    for i, q in enumerate(positions):
        packed |= ((outcomes >> q) & 1) << i
    return packed


class SamplingDistribution(Mapping):
    """Outcome -> probability over the measured qubits.  Keys are LSB-first
    bit tuples; ``outcomes`` holds the packed integers in ascending order."""

    def __init__(self, outcomes: np.ndarray, probabilities: np.ndarray, n_bits: int,
                 n_shots: Optional[int] = None):
        self.outcomes = np.asarray(outcomes, dtype=np.int64)
        self.probabilities = np.asarray(probabilities, dtype=np.float64)
        self.n_bits = int(n_bits)
        self.n_shots = n_shots

    def _tuple(self, v: int) -> tuple:
        return tuple((v >> i) & 1 for i in range(self.n_bits))

    def __getitem__(self, bits) -> float:
        if isinstance(bits, (int, np.integer)):
            v = int(bits)
        else:
            bits = tuple(bits)
            if len(bits) != self.n_bits:
                raise KeyError(bits)
            v = sum(int(b) << i for i, b in enumerate(bits))
        i = int(np.searchsorted(self.outcomes, v))
        if i < self.outcomes.size and self.outcomes[i] == v:
            return float(self.probabilities[i])
        raise KeyError(bits)

    def __iter__(self) -> Iterator[tuple]:
        return (self._tuple(int(v)) for v in self.outcomes)

    def __len__(self) -> int:
        return int(self.outcomes.size)

    def __repr__(self) -> str:
        return f"SamplingDistribution({len(self)} outcomes over {self.n_bits} bits, n_shots={self.n_shots})"


def distribution_from_result(result, measured: Sequence[int]) -> SamplingDistribution:
    """Marginal of ``result`` on ``measured``, normalised by ``n_shots``."""
    outcomes, weights = outcome_arrays(result)
    weights = weights / result.n_shots
    keys, inverse = np.unique(pack_bits(outcomes, measured), return_inverse=True)
    n_shots = None if getattr(result, "is_exact", False) else int(result.n_shots)
    return SamplingDistribution(keys, np.bincount(inverse, weights=weights, minlength=len(keys)),
                                len(list(measured)), n_shots)
