"""Circuit IR consumed by the MI355X backend.

Python counterparts of the qarpx types the simulator boundary takes:
``GateType`` (src/qarpx/include/qarpx/core/gates.h:8-57, same order and
names), ``Param`` (core/param.h:23-105: a concrete double or an affine
``sum_k c_k * symbol_k + offset`` expression, plus a compound escape hatch
that evaluates but is not differentiable), ``Command``
(core/command.h:29-103) with ``dagger`` (core/command.cpp:19-59) and
``substitute``, ``affine_coefficients`` (core/param.cpp:462-497) and
``SamplingResult`` (simulator/sampling_result.h:14-37).
"""
from __future__ import annotations

import enum
import math
from dataclasses import dataclass, field
from typing import Callable, Iterable, Mapping, Optional, Sequence, Union

import numpy as np

from .errors import CapabilityError


class GateType(enum.IntEnum):
    X = 0
    Y = 1
    Z = 2
    H = 3
    S = 4
    Sdg = 5
    T = 6
    Tdg = 7
    SX = 8
    SXdg = 9
    Id = 10
    Rx = 11
    Ry = 12
    Rz = 13
    P = 14
    U = 15
    CX = 16
    CY = 17
    CZ = 18
    SWAP = 19
    ECR = 20
    iSWAP = 21
    iSWAPdg = 22
    CH = 23
    CS = 24
    CSdg = 25
    CSX = 26
    CSXdg = 27
    CRx = 28
    CRy = 29
    CRz = 30
    CP = 31
    RZZ = 32
    RXX = 33
    RYY = 34
    CU = 35
    CCX = 36
    CSWAP = 37
    MCZ = 38
    GPhase = 39
    Barrier = 40
    Measure = 41
    Reset = 42
    BranchBegin = 43
    BranchElse = 44
    BranchEnd = 45
    Custom = 46


_N_PARAMS = {
    GateType.GPhase: 1, GateType.Rx: 1, GateType.Ry: 1, GateType.Rz: 1, GateType.P: 1,
    GateType.CRx: 1, GateType.CRy: 1, GateType.CRz: 1, GateType.CP: 1,
    GateType.RZZ: 1, GateType.RXX: 1, GateType.RYY: 1, GateType.U: 3, GateType.CU: 4,
}

_SELF_ADJOINT = {
    GateType.X, GateType.Y, GateType.Z, GateType.H, GateType.Id, GateType.CX, GateType.CY,
    GateType.CZ, GateType.SWAP, GateType.CH, GateType.CCX, GateType.CSWAP, GateType.MCZ,
    GateType.ECR, GateType.Barrier, GateType.Reset,
}

_NAMED_INVERSE = {
    GateType.S: GateType.Sdg, GateType.Sdg: GateType.S,
    GateType.T: GateType.Tdg, GateType.Tdg: GateType.T,
    GateType.iSWAP: GateType.iSWAPdg, GateType.iSWAPdg: GateType.iSWAP,
    GateType.SX: GateType.SXdg, GateType.SXdg: GateType.SX,
    GateType.CS: GateType.CSdg, GateType.CSdg: GateType.CS,
    GateType.CSX: GateType.CSXdg, GateType.CSXdg: GateType.CSX,
}


def gate_num_params(g: GateType) -> int:
    """gate_num_params (core/gates.h:113-128)."""
    return _N_PARAMS.get(GateType(g), 0)


class Param:
    """A gate parameter: concrete float, affine symbolic expression, or a
    compound expression carried by a closure (evaluable, not affine)."""

    __slots__ = ("_terms", "_offset", "_fn", "_fn_symbols", "_repr")

    def __init__(self, value: Union[float, "Param"] = 0.0):
        if isinstance(value, Param):
            self._terms = dict(value._terms)
            self._offset = value._offset
            self._fn = value._fn
            self._fn_symbols = value._fn_symbols
            self._repr = value._repr
            return
        self._terms: dict[str, float] = {}
        self._offset = float(value)
        self._fn: Optional[Callable[[Mapping[str, float]], float]] = None
        self._fn_symbols: tuple[str, ...] = ()
        self._repr = ""

    # ── constructors ──
    @staticmethod
    def symbol(name: str) -> "Param":
        return Param.linear(1.0, name, 0.0)

    @staticmethod
    def linear(coeff: float, symbol_name: str, offset: float = 0.0) -> "Param":
        p = Param(offset)
        p._terms = {str(symbol_name): float(coeff)}
        return p

    @staticmethod
    def compound(fn: Callable[[Mapping[str, float]], float], symbols: Sequence[str],
                 text: str = "") -> "Param":
        """Escape hatch for expressions outside the affine form."""
        p = Param(0.0)
        p._fn = fn
        p._fn_symbols = tuple(symbols)
        p._repr = text or f"f({', '.join(symbols)})"
        return p

    # ── queries ──
    def is_symbolic(self) -> bool:
        return bool(self._terms) or self._fn is not None

    def is_concrete(self) -> bool:
        return not self.is_symbolic()

    def value(self) -> float:
        if self.is_symbolic():
            raise RuntimeError(f"Param::value on symbolic parameter {self!r}")
        return self._offset

    def free_symbols(self) -> list[str]:
        if self._fn is not None:
            return list(self._fn_symbols)
        return list(self._terms)

    def evaluate(self, values: Mapping[str, float]) -> float:
        missing = [s for s in self.free_symbols() if s not in values]
        if missing:
            raise RuntimeError(f"Param::evaluate: unbound symbol(s) {missing}")
        if self._fn is not None:
            return float(self._fn(values))
        return self._offset + sum(c * float(values[s]) for s, c in self._terms.items())

    def substitute(self, values: Mapping[str, float]) -> "Param":
        """Binds all-or-nothing for compound expressions (conventions §10)."""
        if self._fn is not None:
            if all(s in values for s in self._fn_symbols):
                return Param(self._fn(values))
            return Param(self)
        if not any(s in values for s in self._terms):
            return Param(self)
        out = Param(self._offset)
        for s, c in self._terms.items():
            if s in values:
                out._offset += c * float(values[s])
            else:
                out._terms[s] = c
        return out

    def rename_symbols(self, mapping: Mapping[str, str]) -> "Param":
        if self._fn is not None:
            inv = {mapping.get(s, s): s for s in self._fn_symbols}
            fn = self._fn
            return Param.compound(lambda v: fn({inv.get(k, k): x for k, x in v.items()}),
                                  [mapping.get(s, s) for s in self._fn_symbols], self._repr)
        out = Param(self._offset)
        for s, c in self._terms.items():
            t = mapping.get(s, s)
            out._terms[t] = out._terms.get(t, 0.0) + c
        return out

    # ── arithmetic ──
    @staticmethod
    def _coerce(x) -> "Param":
        return x if isinstance(x, Param) else Param(float(x))

    def _affine(self) -> bool:
        return self._fn is None

    def __neg__(self) -> "Param":
        return self * -1.0

    def __add__(self, other) -> "Param":
        o = Param._coerce(other)
        if self._affine() and o._affine():
            out = Param(self._offset + o._offset)
            out._terms = dict(self._terms)
            for s, c in o._terms.items():
                out._terms[s] = out._terms.get(s, 0.0) + c
            return out
        a, b = Param(self), Param(o)
        syms = list(dict.fromkeys(a.free_symbols() + b.free_symbols()))
        return Param.compound(lambda v: a.evaluate(v) + b.evaluate(v), syms, f"({a!r} + {b!r})")

    __radd__ = __add__

    def __sub__(self, other) -> "Param":
        return self + (-Param._coerce(other))

    def __rsub__(self, other) -> "Param":
        return Param._coerce(other) - self

    def __mul__(self, other) -> "Param":
        o = Param._coerce(other)
        if o.is_concrete() and self._affine():
            k = o._offset
            out = Param(self._offset * k)
            out._terms = {s: c * k for s, c in self._terms.items()}
            return out
        if self.is_concrete() and o._affine():
            return o * self._offset
        a, b = Param(self), Param(o)
        syms = list(dict.fromkeys(a.free_symbols() + b.free_symbols()))
        return Param.compound(lambda v: a.evaluate(v) * b.evaluate(v), syms, f"({a!r} * {b!r})")

    __rmul__ = __mul__

    def __truediv__(self, other) -> "Param":
        o = Param._coerce(other)
        if o.is_concrete():
            return self * (1.0 / o._offset)
        a, b = Param(self), Param(o)
        syms = list(dict.fromkeys(a.free_symbols() + b.free_symbols()))
        return Param.compound(lambda v: a.evaluate(v) / b.evaluate(v), syms, f"({a!r} / {b!r})")

    def __eq__(self, other) -> bool:
        if not isinstance(other, Param):
            try:
                other = Param(float(other))
            except (TypeError, ValueError):
                return NotImplemented
        if self._fn is not None or other._fn is not None:
            return self is other
        return self._offset == other._offset and self._terms == other._terms

    def __hash__(self):
        return hash((self._offset, tuple(sorted(self._terms.items()))))

    def approx_equal(self, other, atol: float = 1e-9) -> bool:
        o = Param._coerce(other)
        if self._fn is not None or o._fn is not None:
            return repr(self) == repr(o)
        if set(self._terms) != set(o._terms):
            return False
        return abs(self._offset - o._offset) <= atol and all(
            abs(c - o._terms[s]) <= atol for s, c in self._terms.items())

    def __float__(self) -> float:
        return self.value()

    def __repr__(self) -> str:
        if self._fn is not None:
            return self._repr
        if not self._terms:
            return repr(self._offset)
        parts = [f"{c:g}*{s}" for s, c in self._terms.items()]
        if self._offset:
            parts.append(f"{self._offset:g}")
        return " + ".join(parts)


def symbol(name: str) -> Param:
    return Param.symbol(name)


def concrete(value: float) -> Param:
    return Param(value)


def affine_coefficients(param) -> dict[str, float]:
    """d(param)/d(symbol) for every free symbol (core/param.cpp:462-497).

    Affine parameters of this package answer with their coefficients;
    compound ones and the reference's own Param objects are probed as the
    reference does: at 0, then 1 and 2 per symbol (coefficient = f(1) - f(0)),
    then all-ones, each check within 1e-9 of the running scale; an
    evaluation that fails or is not finite, or a failed check, raises
    CapabilityError."""
    p = as_param(param)
    if p.is_concrete():
        return {}
    if p._fn is None:
        return dict(p._terms)

    def probe(at):
        try:
            v = p.evaluate(at)
        except (RuntimeError, ZeroDivisionError, ValueError, OverflowError) as e:
            raise CapabilityError(f"parameter {p!r} cannot be evaluated at a probe point ({e}): gradient "
                                  "rules are exact only for affine gate angles") from None
        if not math.isfinite(v):
            raise CapabilityError(f"parameter {p!r} evaluates to a non-finite value at a probe point")
        return v

    syms = p.free_symbols()
    at = {s: 0.0 for s in syms}
    v0 = probe(at)
    scale = max(1.0, abs(v0))
    coeffs, total = {}, 0.0
    for s in syms:
        at[s] = 1.0
        v1 = probe(at)
        at[s] = 2.0
        v2 = probe(at)
        at[s] = 0.0
        c = v1 - v0
        scale = max(scale, abs(v1), abs(v2))
        if abs((v2 - v0) - 2.0 * c) > 1e-9 * scale:
            raise CapabilityError(f"parameter {p!r} enters non-linearly in symbol '{s}': gradient rules are "
                                  "exact only for affine gate angles")
        coeffs[s] = c
        total += c
    for s in syms:
        at[s] = 1.0
    v_all = probe(at)
    if abs((v_all - v0) - total) > 1e-9 * max(scale, abs(v_all)):
        raise CapabilityError(f"parameter {p!r} is not affine in its symbols: gradient rules are exact only "
                              "for affine gate angles")
    return coeffs


ParamLike = Union[Param, float, int]


@dataclass
class Command:
    """One gate operation (core/command.h:29-103)."""

    gate: GateType
    qubits: tuple = ()
    params: tuple = ()
    cbits: tuple = ()
    condition_bits: tuple = ()
    condition_values: tuple = ()
    unitary: Optional[np.ndarray] = field(default=None, repr=False, compare=False)

    def __init__(self, gate: GateType = GateType.Barrier, *args, qubits: Iterable[int] = (),
                 params: Iterable[ParamLike] = (), cbits: Iterable[int] = (),
                 condition_bits: Iterable[int] = (), condition_values: Iterable[bool] = (),
                 unitary: Optional[np.ndarray] = None):
        qs, ps = list(qubits), [Param._coerce(p) for p in params]
        # Positional form of the reference constructors: Command(g, q),
        # Command(g, q0, q1), Command(g, q, param), Command(g, q0, q1, param).
        positional_param = False
        for a in args:
            if isinstance(a, (Param, float, np.floating)):
                ps.append(Param._coerce(a))
                positional_param = True
            elif isinstance(a, (int, np.integer)) and not positional_param:
                qs.append(int(a))
            else:
                raise TypeError(f"Command: unexpected argument {a!r}")
        self.gate = GateType(gate)
        self.qubits = tuple(int(q) for q in qs)
        self.params = tuple(ps)
        self.cbits = tuple(int(c) for c in cbits)
        self.condition_bits = tuple(int(c) for c in condition_bits)
        self.condition_values = tuple(bool(v) for v in condition_values)
        if unitary is not None:
            unitary = np.ascontiguousarray(unitary, dtype=np.complex128)
            d = 1 << len(self.qubits)
            if unitary.shape != (d, d):
                raise ValueError(f"Custom unitary is {unitary.shape} for {len(self.qubits)} qubit(s)")
        self.unitary = unitary

    def is_parametric(self) -> bool:
        return any(p.is_symbolic() for p in self.params)

    def dagger(self) -> "Command":
        """Command::dagger (core/command.cpp:19-59)."""
        out = self.copy()
        g = self.gate
        if g in _SELF_ADJOINT:
            return out
        if g in _NAMED_INVERSE:
            out.gate = _NAMED_INVERSE[g]
            return out
        if g == GateType.U:
            t, ph, la = self.params
            out.params = (-t, -la, -ph)
            return out
        if g == GateType.CU:
            t, ph, la, ga = self.params
            out.params = (-t, -la, -ph, -ga)
            return out
        if g == GateType.Custom and self.unitary is not None:
            out.unitary = np.ascontiguousarray(self.unitary.conj().T)
            return out
        out.params = tuple(-p for p in self.params)
        return out

    def copy(self) -> "Command":
        return Command(self.gate, qubits=self.qubits, params=self.params, cbits=self.cbits,
                       condition_bits=self.condition_bits, condition_values=self.condition_values,
                       unitary=self.unitary)

    def substitute(self, values: Mapping[str, float]) -> "Command":
        out = self.copy()
        out.params = tuple(p.substitute(values) for p in self.params)
        return out

    def remap_qubits(self, mapping: Sequence[int]) -> "Command":
        out = self.copy()
        out.qubits = tuple(int(mapping[q]) for q in self.qubits)
        return out

    def rename_symbols(self, mapping: Mapping[str, str]) -> "Command":
        out = self.copy()
        out.params = tuple(p.rename_symbols(mapping) for p in self.params)
        return out

    def approx_equal(self, other: "Command", atol: float = 1e-9) -> bool:
        return (self.gate == other.gate and self.qubits == other.qubits
                and self.cbits == other.cbits and self.condition_bits == other.condition_bits
                and self.condition_values == other.condition_values
                and len(self.params) == len(other.params)
                and all(a.approx_equal(b, atol) for a, b in zip(self.params, other.params)))

    def __eq__(self, other) -> bool:
        if not isinstance(other, Command):
            return NotImplemented
        return (self.gate == other.gate and self.qubits == other.qubits
                and self.params == other.params and self.cbits == other.cbits
                and self.condition_bits == other.condition_bits
                and self.condition_values == other.condition_values)

    def __repr__(self) -> str:
        ps = f", params={list(self.params)}" if self.params else ""
        return f"Command({self.gate.name}, qubits={list(self.qubits)}{ps})"


# ── the reference's own objects at the boundary ────────────────────────────
# QarpEngine hands the simulator qx.Command objects (src/qarpx/python/
# qarpx.pyi:576-620: gate, qubits, params, cbits, condition_bits,
# condition_values) whose params are qx.Param (:507-555: is_concrete, value,
# free_symbols, evaluate).  They are read through that public surface only.

def as_param(p) -> Param:
    """A parameter of this package from a float, a Param, or an object with
    qx.Param's surface.  A symbolic foreign parameter keeps its own evaluate()
    (so bound values equal the reference's) and answers affine_coefficients
    by the reference's probes (core/param.cpp:462-497)."""
    if isinstance(p, Param):
        return p
    if isinstance(p, (float, int, np.floating, np.integer)):
        return Param(float(p))
    is_concrete = getattr(p, "is_concrete", None)
    if is_concrete is None:
        return Param(float(p))
    if is_concrete():
        return Param(float(p.value()))
    evaluate = p.evaluate
    return Param.compound(lambda values: float(evaluate(dict(values))), list(p.free_symbols()), repr(p))


def _gate_of(g) -> GateType:
    if isinstance(g, GateType):
        return g
    v = getattr(g, "value", g)
    try:
        return GateType(int(v))
    except (TypeError, ValueError):
        return GateType[str(getattr(g, "name", g))]


def as_command(cmd) -> Command:
    """A Command of this package from one of ours or an object with
    qx.Command's surface (a ``unitary`` attribute is read when present)."""
    if isinstance(cmd, Command):
        return cmd
    return Command(_gate_of(cmd.gate), qubits=list(cmd.qubits), params=[as_param(p) for p in cmd.params],
                   cbits=list(getattr(cmd, "cbits", ()) or ()),
                   condition_bits=list(getattr(cmd, "condition_bits", ()) or ()),
                   condition_values=list(getattr(cmd, "condition_values", ()) or ()),
                   unitary=getattr(cmd, "unitary", None))


def as_commands(commands) -> list:
    """``commands`` as Commands of this package; a list of them is returned
    as it is."""
    cmds = commands if isinstance(commands, list) else list(commands)
    for c in cmds:
        if not isinstance(c, Command):
            return [as_command(x) for x in cmds]
    return cmds


def substitute_all(commands: Sequence[Command], values: Mapping[str, float]) -> list[Command]:
    return [c.substitute(values) for c in commands]


def make_gphase(param: ParamLike) -> Command:
    return Command(GateType.GPhase, params=[param])


@dataclass
class SamplingResult:
    """Outcome counts of one execution (simulator/sampling_result.h:14-37)."""

    n_qubits: int = 0
    n_shots: int = 0
    counts: dict = field(default_factory=dict)
    n_cbits: int = 0
    cbit_history: list = field(default_factory=list)
    statevector: list = field(default_factory=list)

    def probability(self, outcome: int) -> float:
        c = self.counts.get(int(outcome), 0)
        return c / self.n_shots if self.n_shots else 0.0


def cbit_register_width(commands: Sequence[Command]) -> int:
    """cbit_register_width (core/command.h:185-191)."""
    w = 0
    for c in commands:
        if c.gate == GateType.Measure:
            for b in c.cbits:
                w = max(w, b + 1)
        for b in c.condition_bits:
            w = max(w, b + 1)
    return w


PI = math.pi
