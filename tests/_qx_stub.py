"""Stand-ins with the Python surface of the reference's qarpx types, as its
stub declares them (src/qarpx/python/qarpx.pyi:399-487 GateType, :507-555
Param, :576-620 Command): what QarpEngine hands the simulator.  Test
infrastructure only: they let the drop-in be exercised without the
reference's extension module (which needs Python >= 3.11 and nanobind).

GateType is a plain enum.Enum (not an int), Param is ``coeff * symbol +
offset`` or a compound closure, evaluated as the reference does
(core/param.cpp: substitute returns coeff * value + offset), and Command has
no ``unitary`` attribute.
"""
from __future__ import annotations

import enum

from openqarp_amd.ir import GateType as _Ours

GateType = enum.Enum("GateType", [(g.name, g.value) for g in _Ours])


class Param:
    def __init__(self, value: float = 0.0):
        self._v = float(value)
        self._sym = None
        self._coeff = 1.0
        self._off = 0.0
        self._fn = None
        self._syms: list = []

    @staticmethod
    def symbol(name: str) -> "Param":
        return Param.linear(1.0, name, 0.0)

    @staticmethod
    def linear(coeff: float, symbol_name: str, offset: float = 0.0) -> "Param":
        p = Param()
        p._sym, p._coeff, p._off = symbol_name, float(coeff), float(offset)
        return p

    @staticmethod
    def compound(fn, symbols) -> "Param":
        p = Param()
        p._fn, p._syms = fn, list(symbols)
        return p

    def is_symbolic(self) -> bool:
        return self._sym is not None or self._fn is not None

    def is_concrete(self) -> bool:
        return not self.is_symbolic()

    def value(self) -> float:
        if self.is_symbolic():
            raise RuntimeError("Param::value() on a symbolic parameter")
        return self._v

    def free_symbols(self) -> list:
        if self._fn is not None:
            return list(self._syms)
        return [] if self._sym is None else [self._sym]

    def evaluate(self, values) -> float:
        if self._fn is not None:
            missing = [s for s in self._syms if s not in values]
            if missing:
                raise RuntimeError(f"Param::evaluate() - unresolved symbols {missing}")
            return float(self._fn(values))
        if self._sym is None:
            return self._v
        if self._sym not in values:
            raise RuntimeError(f"Param::evaluate() - unresolved symbol {self._sym}")
        return self._coeff * float(values[self._sym]) + self._off

    def __repr__(self) -> str:
        if self._fn is not None:
            return f"f({', '.join(self._syms)})"
        return repr(self._v) if self._sym is None else f"{self._coeff}*{self._sym} + {self._off}"


class Command:
    def __init__(self, gate=None, *args, params=None, cbits=(), condition_bits=(), condition_values=()):
        self.gate = gate if gate is not None else GateType.Barrier
        qs, ps = [], []
        for a in args:
            if isinstance(a, (Param, float)):
                ps.append(a if isinstance(a, Param) else Param(a))
            else:
                qs.append(int(a))
        self._qubits = qs
        self._params = ps + [p if isinstance(p, Param) else Param(p) for p in (params or [])]
        self._cbits = list(cbits)
        self._cond_bits = list(condition_bits)
        self._cond_values = list(condition_values)

    @property
    def qubits(self) -> list:
        return list(self._qubits)

    @property
    def params(self) -> list:
        return list(self._params)

    @property
    def cbits(self) -> list:
        return list(self._cbits)

    @property
    def condition_bits(self) -> list:
        return list(self._cond_bits)

    @property
    def condition_values(self) -> list:
        return list(self._cond_values)

    def is_parametric(self) -> bool:
        return any(p.is_symbolic() for p in self._params)


def from_ours(cmd) -> Command:
    """The stand-in of one of this package's Commands (no Custom: the
    reference's Python surface cannot build one)."""
    out = Command(GateType[cmd.gate.name], *cmd.qubits, cbits=cmd.cbits, condition_bits=cmd.condition_bits,
                  condition_values=cmd.condition_values)
    for p in cmd.params:
        if p.is_concrete():
            out._params.append(Param(p.value()))
        else:
            ((s, c),) = p._terms.items()
            out._params.append(Param.linear(c, s, p._offset))
    return out
