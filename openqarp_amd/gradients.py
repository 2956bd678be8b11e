"""Gradient methods of the engine: adjoint (one device sweep), parameter shift
and finite differences (stencils of device forward + ⟨H⟩ evaluations).

* Shift rules (src/qarp/engines/_gradients.py:250-301), in angle units as
  (shift, coefficient) lists: the two-term rule (±π/2, ±½) for the Pauli
  rotations Rx, Ry, Rz, RXX, RYY, RZZ and the projector phases P, CP; the
  four-term rule of Anselmetti et al. (frequencies ½ and 1: ±π/2 with
  ±(√2+1)/(4√2), ±3π/2 with ∓(√2−1)/(4√2)) for CRx, CRy, CRz; nothing for
  GPhase, which drops out of ⟨ψ|H|ψ⟩.
* Every occurrence of a symbol is shifted on its own (ShiftPlan,
  :317-367): occurrences get private symbol names, the affine coefficient c
  of angle = c·x + b scales the shift (s/c in symbol units) and the
  coefficient (c·w).
* A symbolic U(θ, φ, λ) is rewritten first as Rz(λ), Ry(θ), Rz(φ) and
  GPhase((φ+λ)/2) (the phase-exact decomposition of
  rebase_differentiable, :372-381); a symbolic CU has no rewrite here and
  raises CapabilityError.
* Finite differences (:427-443): central (x ± ε e_k, ±1/(2ε)) or forward,
  ε = 1e-5 by default.
* Adjoint (:495-546): run_gradient for an expectation value; for an overlap,
  ∂|⟨bra|ket⟩|² = 2 Re(⟨bra|ket⟩* ⟨bra|∂ket⟩) from run_gradient_phi with
  |φ⟩ = ⟨bra|ket⟩|bra⟩, plus the mirrored sweep over the bra when it carries
  a tracked symbol.
"""
from __future__ import annotations

import math
from typing import Mapping, Optional, Sequence

import numpy as np

from .errors import CapabilityError
from .ir import Command, GateType, Param, affine_coefficients

GRADIENT_METHODS = ("default", "adjoint", "parameter-shift", "finite-diff")

_TWO_TERM_HALF = [(math.pi / 2, 0.5), (-math.pi / 2, -0.5)]
_C_PLUS = (math.sqrt(2.0) + 1.0) / (4.0 * math.sqrt(2.0))
_C_MINUS = (math.sqrt(2.0) - 1.0) / (4.0 * math.sqrt(2.0))
_FOUR_TERM = [(math.pi / 2, _C_PLUS), (-math.pi / 2, -_C_PLUS),
              (3 * math.pi / 2, -_C_MINUS), (-3 * math.pi / 2, _C_MINUS)]

SHIFT_RULES = {
    **{g: _TWO_TERM_HALF for g in (GateType.Rx, GateType.Ry, GateType.Rz, GateType.RXX, GateType.RYY,
                                   GateType.RZZ, GateType.P, GateType.CP)},
    **{g: _FOUR_TERM for g in (GateType.CRx, GateType.CRy, GateType.CRz)},
    GateType.GPhase: [],
}


def shift_rule(gate: GateType) -> list:
    try:
        return SHIFT_RULES[GateType(gate)]
    except KeyError:
        raise CapabilityError(
            f"gate '{GateType(gate).name}' with a symbolic angle has no parameter-shift rule "
            "(symbolic U is rewritten before differentiation; anything else is unsupported).  "
            "Use method='finite-diff'.") from None


def rebase_differentiable(commands: Sequence[Command]) -> list:
    """Symbolic U(θ, φ, λ) -> Rz(λ) Ry(θ) Rz(φ) GPhase((φ+λ)/2), applied in that
    order (U = e^{i(φ+λ)/2} Rz(φ) Ry(θ) Rz(λ)); other commands pass through."""
    out = []
    for c in commands:
        if c.gate == GateType.U and c.is_parametric():
            (q,) = c.qubits
            t, ph, la = c.params
            out += [Command(GateType.Rz, q, la), Command(GateType.Ry, q, t), Command(GateType.Rz, q, ph),
                    Command(GateType.GPhase, params=[(ph + la) * 0.5])]
        elif c.gate == GateType.CU and c.is_parametric():
            raise CapabilityError("a symbolic CU has no differentiable rewrite in this backend; "
                                  "use method='finite-diff' or the adjoint method")
        else:
            out.append(c)
    return out


class ShiftPlan:
    """Commands with one private symbol per occurrence of each tracked symbol,
    and the stencil of every tracked symbol."""

    def __init__(self, commands: Sequence[Command], symbols: Sequence[str]):
        tracked = set(symbols)
        self.commands: list = []
        self.occurrences: dict = {s: [] for s in symbols}   # sym -> [(private, coeff, gate)]
        self._inert: list = []                               # (sym, private) with coefficient 0
        n = 0
        for c in rebase_differentiable(commands):
            if not (c.is_parametric() and tracked.intersection(s for p in c.params for s in p.free_symbols())):
                self.commands.append(c)
                continue
            params = []
            for p in c.params:
                mapping = {}
                for s in p.free_symbols():
                    if s in tracked:
                        private = f"{s}\0{n}"
                        n += 1
                        mapping[s] = private
                q = p.rename_symbols(mapping) if mapping else p
                coeffs = affine_coefficients(q) if mapping else {}
                for s, private in mapping.items():
                    if coeffs.get(private, 0.0) == 0.0:
                        self._inert.append((s, private))
                    else:
                        self.occurrences[s].append((private, coeffs[private], c.gate))
                params.append(q)
            d = c.copy()
            d.params = tuple(params)
            self.commands.append(d)

    def base_point(self, params: Mapping[str, float]) -> dict:
        base = dict(params)
        for s, occs in self.occurrences.items():
            for private, _, _ in occs:
                base[private] = params[s]
        for s, private in self._inert:
            base[private] = params[s]
        return base

    def stencil(self, sym: str) -> list:
        out = []
        for private, coeff, gate in self.occurrences[sym]:
            for shift, w in shift_rule(gate):
                out.append(({private: shift / coeff}, w * coeff))
        return out


def finite_difference_stencils(symbols: Sequence[str], fd_eps: float = 1e-5, fd_order: int = 2) -> list:
    if fd_order not in (1, 2):
        raise ValueError(f"fd_order must be 1 (forward) or 2 (central), got {fd_order!r}")
    eps = float(fd_eps)
    if eps <= 0.0:
        raise ValueError(f"fd_eps must be positive, got {fd_eps!r}")
    if fd_order == 1:
        return [[({}, -1.0 / eps), ({s: eps}, 1.0 / eps)] for s in symbols]
    return [[({s: eps}, 0.5 / eps), ({s: -eps}, -0.5 / eps)] for s in symbols]


def evaluate_stencils(evaluate, base: Mapping[str, float], stencils: Sequence[list]) -> np.ndarray:
    """Σ_points w·f(base + offset) per symbol; all distinct points go to
    ``evaluate(list_of_param_dicts) -> values`` in one batch."""
    points: dict = {}
    order = []
    for st in stencils:
        for offset, _ in st:
            key = tuple(sorted(offset.items()))
            if key not in points:
                points[key] = len(order)
                p = dict(base)
                for s, d in offset.items():
                    p[s] = p[s] + d
                order.append(p)
    values = np.asarray(evaluate(order), dtype=np.float64) if order else np.zeros(0)
    out = np.zeros(len(stencils))
    for k, st in enumerate(stencils):
        out[k] = sum(w * values[points[tuple(sorted(off.items()))]] for off, w in st)
    return out


def _with_ket(prim, commands):
    """The primitive with its (ket) circuit replaced."""
    from .primitives import PauliAveraging, StateVector
    if isinstance(prim, StateVector):
        return StateVector(commands, prim.n_qubits, prim.observable, bra=prim.bra, target=prim.target,
                           initial_state=prim.initial_state)
    if isinstance(prim, PauliAveraging):
        return PauliAveraging(commands, prim.n_qubits, prim.observable, prim.n_shots, prim.grouping)
    raise CapabilityError(f"{type(prim).__name__} has no stencil gradient")


# Device memory one batched evaluation may hold in states (psi and H psi of
# every point of a chunk).
BATCH_STATE_BYTES = 32 << 30


def _point_evaluator(sim, prim):
    """Batched f(params) for a primitive.  An expectation value from |0...0>
    goes through qhip_batch_energy_grad (energies only), one batched launch
    per schedule step for a chunk of points; anything else runs point by
    point."""
    from .primitives import StateVector, Target
    if isinstance(prim, StateVector) and prim.target is Target.EXPECTATION_VALUE and prim.initial_state is None:
        amp = 16 if sim.precision == "complex128" else 8
        per = max(1, BATCH_STATE_BYTES // (2 * amp << prim.n_qubits))

        def run(pts):
            out = []
            for k in range(0, len(pts), per):
                chunk = [(prim.ket, prim.n_qubits, prim.observable, p, []) for p in pts[k:k + per]]
                e, _ = sim.batch_energy_grad(chunk, gradients=False)
                out.extend(e.tolist())
            return out
        return run
    return lambda pts: [float(prim.run(sim, p)) for p in pts]


def _check_kind(prim, method: str) -> None:
    if getattr(prim, "gradient_kind", "none") != "expectation":
        raise CapabilityError(f"{type(prim).__name__}: method '{method}' needs a real expectation value "
                              f"(gradient_kind 'expectation'), got '{getattr(prim, 'gradient_kind', 'none')}'")


def parameter_shift_gradient(sim, prim, params: Mapping[str, float], symbols: Sequence[str]) -> np.ndarray:
    _check_kind(prim, "parameter-shift")
    plan = ShiftPlan(prim.ket, symbols)
    shifted = _with_ket(prim, plan.commands)
    return evaluate_stencils(_point_evaluator(sim, shifted), plan.base_point(params),
                             [plan.stencil(s) for s in symbols])


def finite_diff_gradient(sim, prim, params: Mapping[str, float], symbols: Sequence[str], *,
                         fd_eps: float = 1e-5, fd_order: int = 2) -> np.ndarray:
    _check_kind(prim, "finite-diff")
    return evaluate_stencils(_point_evaluator(sim, prim), dict(params),
                             finite_difference_stencils(symbols, fd_eps, fd_order))


def adjoint_gradient(sim, prim, params: Mapping[str, float], symbols: Sequence[str]) -> np.ndarray:
    from .primitives import StateVector, Target, bind
    if not isinstance(prim, StateVector):
        raise CapabilityError(f"{type(prim).__name__}: the adjoint method needs amplitudes (StateVector)")
    ket_cmds = rebase_differentiable(prim.ket)
    if prim.target is Target.EXPECTATION_VALUE:
        return np.asarray(sim.run_gradient(ket_cmds, prim.n_qubits, prim.observable, params, list(symbols),
                                           initial_state=prim.initial_state))
    if prim.target is not Target.OVERLAP:
        raise CapabilityError("the adjoint method differentiates EXPECTATION_VALUE and OVERLAP")
    ket = sim.statevector(bind(prim.ket, params), prim.n_qubits, initial_state=prim.initial_state)
    bra = sim.statevector(bind(prim.bra, params), prim.n_qubits)
    ov = complex(np.vdot(bra, ket))
    grad = np.asarray(sim.run_gradient_phi(ket_cmds, prim.n_qubits, ov * bra, params, list(symbols),
                                           initial_state=prim.initial_state))
    tracked = set(symbols)
    if any(tracked.intersection(s for p in c.params for s in p.free_symbols()) for c in prim.bra):
        grad = grad + np.asarray(sim.run_gradient_phi(rebase_differentiable(prim.bra), prim.n_qubits,
                                                      np.conj(ov) * ket, params, list(symbols)))
    return grad
