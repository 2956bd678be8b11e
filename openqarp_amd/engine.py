"""The engine: the reference's Engine surface (src/qarp/engines/_engine.py,
_qarp_engine.py) reduced to what drives the statevector path — run a
primitive, sweep parameter sets, differentiate.

``HipEngine`` owns one :class:`QarpSimulator` (the object QarpEngine keeps in
``self._sim``, _qarp_engine.py:198-203).  Sampling seeds follow the
reference: circuit ordinal k of a call draws with (seed + 100003·k) mod 2^32
(_engine.py:383-394) and ``batch_run`` offsets parameter set i by +i
(qarp_simulator.cpp:1261-1263).  ``run_gradient`` dispatches on ``method``
as Engine.run_gradient does (_gradients.py:29-35); "default" is the adjoint
sweep where the primitive has amplitudes, parameter shift otherwise.
"""
from __future__ import annotations

from typing import Mapping, Optional, Sequence

import numpy as np

from . import gradients as G
from .errors import CapabilityError
from .primitives import PauliAveraging, Sampler, StateVector
from .simulator import QarpSimulator


class HipEngine:
    gradient_methods = frozenset(G.GRADIENT_METHODS)

    def __init__(self, *, n_shots: Optional[int] = None, seed: Optional[int] = None,
                 precision: str = "complex128", fusion_max_qubits: Optional[int] = None,
                 fusion_min_qubits: Optional[int] = None, simulator: Optional[QarpSimulator] = None):
        self.sim = simulator if simulator is not None else QarpSimulator(precision=precision)
        if fusion_max_qubits is not None:
            self.sim.fusion_max_qubits = fusion_max_qubits
        if fusion_min_qubits is not None:
            self.sim.fusion_min_qubits = fusion_min_qubits
        self._n_shots = n_shots
        self._seed = None if seed is None else int(seed) & 0xFFFFFFFF

    def _circuit_seed(self, ordinal: int) -> Optional[int]:
        if self._seed is None:
            return None
        return (self._seed + 100_003 * ordinal) % 2**32

    def run(self, prim, params: Optional[Mapping[str, float]] = None):
        """Result of one primitive: a float / complex (StateVector,
        PauliAveraging) or a SamplingDistribution (Sampler)."""
        params = dict(params or {})
        if isinstance(prim, Sampler):
            return prim.run(self.sim, params, self._circuit_seed(0))
        if isinstance(prim, PauliAveraging):
            return prim.run(self.sim, params, [self._circuit_seed(k) for k in range(prim.n_groups)])
        if isinstance(prim, StateVector):
            return prim.run(self.sim, params)
        raise CapabilityError(f"HipEngine cannot run {type(prim).__name__}")

    def batch_run(self, prim, param_sets: Sequence[Mapping[str, float]]) -> list:
        if isinstance(prim, Sampler):
            seed0 = self._circuit_seed(0)
            return [prim.run(self.sim, p, None if seed0 is None else (seed0 + i) % 2**32)
                    for i, p in enumerate(param_sets)]
        return [self.run(prim, p) for p in param_sets]

    def run_gradient(self, prim, params: Mapping[str, float], symbols: Optional[Sequence[str]] = None, *,
                     method: str = "default", **options) -> np.ndarray:
        """∂ result / ∂ symbol for every symbol in ``symbols`` (default: the
        primitive's free symbols, in order of appearance)."""
        if method not in G.GRADIENT_METHODS:
            raise ValueError(f"unknown gradient method {method!r}; expected one of {sorted(G.GRADIENT_METHODS)}")
        symbols = list(prim.symbols() if symbols is None else symbols)
        params = dict(params)
        missing = [s for s in symbols if s not in params]
        if missing:
            raise ValueError(f"run_gradient: no value for symbols {missing}")
        if method == "default":
            method = "adjoint" if isinstance(prim, StateVector) else "parameter-shift"
        if method == "adjoint":
            return G.adjoint_gradient(self.sim, prim, params, symbols)
        if method == "parameter-shift":
            return G.parameter_shift_gradient(self.sim, prim, params, symbols)
        return G.finite_diff_gradient(self.sim, prim, params, symbols, **options)
