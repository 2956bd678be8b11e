"""The drop-in boundary on the reference's own objects: streams of objects
with qx.Command / qx.Param's surface (tests/_qx_stub.py, shaped like
src/qarpx/python/qarpx.pyi) pack into exactly the records this package's
Commands do, the dependency pools carry the reference's probed affine
coefficients (core/param.cpp:462-497), and the primitives, programs and
scheduler accept them.  No GPU calls."""
import math

import numpy as np
import pytest

import _oracle as O
import _qx_stub as qx
from openqarp_amd import workloads as W
from openqarp_amd._abi import pack_stream
from openqarp_amd.errors import CapabilityError
from openqarp_amd.ir import Command, GateType, Param, affine_coefficients, as_command, as_commands
from openqarp_amd.primitives import PauliAveraging, Sampler, StateVector
from openqarp_amd.program import Program
from openqarp_amd.simulator import inspect_schedule


def _same_records(a, b):
    assert np.array_equal(a.ops, b.ops)
    assert np.array_equal(a.qubits, b.qubits)
    assert np.array_equal(a.dep_index, b.dep_index)
    assert np.array_equal(a.dep_alpha, b.dep_alpha)


def _mixed_circuit(n, seed):
    """Every named gate kind with concrete and symbolic (coeff * s + 0)
    angles, measurements with cbits and a condition."""
    from test_gpu_parity import random_circuit
    cmds = random_circuit(n, 60, seed, with_custom=False)
    rng = np.random.default_rng(seed)
    for q in range(n):
        cmds.append(Command(GateType.Ry, q, Param.linear(float(rng.uniform(0.5, 2)), f"a{q % 3}")))
        cmds.append(Command(GateType.CRz, q, (q + 1) % n, Param.symbol(f"b{q % 2}")))
    cmds.append(Command(GateType.Measure, 0, cbits=[0]))
    cmds.append(Command(GateType.X, 1, condition_bits=[0], condition_values=[True]))
    return cmds


def test_reference_shaped_commands_pack_identically():
    n = 5
    ours = _mixed_circuit(n, 3)
    theirs = [qx.from_ours(c) for c in ours]
    assert all(not isinstance(c, Command) for c in theirs)
    values = {"a0": 0.3, "a1": -1.1, "a2": 2.5, "b0": 0.7, "b1": -0.2}
    index = {s: i for i, s in enumerate(sorted(values))}
    _same_records(pack_stream(as_commands(theirs), values, index), pack_stream(ours, values, index))
    conv = as_commands(theirs)
    for a, b in zip(conv, ours):
        assert (a.gate, a.qubits, a.cbits, a.condition_bits, a.condition_values) == \
               (b.gate, b.qubits, b.cbits, b.condition_bits, b.condition_values)
    assert as_commands(ours) is ours or as_commands(ours) == ours


def test_affine_coefficients_are_the_references_probes():
    """alpha = (c * 1 + o) - (c * 0 + o), as affine_coefficients computes it,
    so a nonzero offset can move alpha by an ulp from c."""
    p = qx.Param.linear(0.1, "t", 0.7)
    alpha = affine_coefficients(p)["t"]
    assert alpha == (0.1 * 1.0 + 0.7) - (0.1 * 0.0 + 0.7)
    assert alpha != 0.1                        # the probe's rounding, kept
    assert affine_coefficients(qx.Param.linear(-2.5, "u", 0.0)) == {"u": -2.5}
    assert affine_coefficients(qx.Param(1.25)) == {}
    two = qx.Param.compound(lambda v: 3.0 * v["x"] - 0.5 * v["y"] + 1.0, ["x", "y"])
    assert affine_coefficients(two) == pytest.approx({"x": 3.0, "y": -0.5}, abs=1e-15)
    with pytest.raises(CapabilityError):
        affine_coefficients(qx.Param.compound(lambda v: math.sin(v["x"]), ["x"]))
    with pytest.raises(CapabilityError):   # a product of two symbols: additive probe fails
        affine_coefficients(qx.Param.compound(lambda v: v["x"] * v["y"], ["x", "y"]))


def test_dependency_pools_of_foreign_params():
    cmds = [qx.Command(qx.GateType.Ry, 0, qx.Param.linear(0.1, "t", 0.7)),
            qx.Command(qx.GateType.RZZ, 0, 1, qx.Param.compound(lambda v: 2.0 * v["t"] - v["s"], ["t", "s"]))]
    ps = pack_stream(as_commands(cmds), {"t": 0.4, "s": -0.3}, {"t": 0, "s": 1})
    assert ps.ops["params"][0][0] == 0.1 * 0.4 + 0.7
    assert ps.ops["params"][1][0] == 2.0 * 0.4 - (-0.3)
    assert list(ps.dep_index) == [0, 0, 1]
    assert ps.dep_alpha[0] == (0.1 + 0.7) - 0.7
    assert list(ps.dep_alpha[1:]) == pytest.approx([2.0, -1.0], abs=1e-15)
    bad = [qx.Command(qx.GateType.Rx, 0, qx.Param.compound(lambda v: v["t"] ** 2, ["t"]))]
    with pytest.raises(CapabilityError, match="run_gradient"):
        pack_stream(as_commands(bad), {"t": 0.4}, {"t": 0})


def test_unbound_foreign_symbol_raises_as_the_reference_does():
    cmds = [qx.Command(qx.GateType.Rx, 0, qx.Param.symbol("t"))]
    with pytest.raises(RuntimeError):
        pack_stream(as_commands(cmds))


def test_oracle_state_of_foreign_stream_equals_native():
    n = 6
    ours = [c for c in _mixed_circuit(n, 9) if c.gate != GateType.Measure and not c.condition_bits]
    theirs = [qx.from_ours(c) for c in ours]
    values = {"a0": 0.3, "a1": -1.1, "a2": 2.5, "b0": 0.7, "b1": -0.2}
    a = O.statevector(pack_stream(as_commands(theirs), values), n, 3, 0)
    b = O.statevector(pack_stream(ours, values), n, 3, 0)
    assert np.array_equal(a, b)


def test_primitives_programs_and_scheduler_accept_foreign_commands():
    n = 8
    w = W.c3_gradient(n, 2, 5, 20)
    theirs = [qx.from_ours(c) for c in w.commands]
    sv = StateVector(theirs, n, w.observable)
    assert sv.symbols() == StateVector(w.commands, n, w.observable).symbols()
    assert all(isinstance(c, Command) for c in sv.ket)
    assert Sampler(theirs, n, 100).symbols() == sv.symbols()
    assert PauliAveraging(theirs, n, w.observable, 100).symbols() == sv.symbols()
    prog = Program()
    prog.add_gates([qx.from_ours(c) for c in W.brickwork_haar(n, 1, 1)[:0]] +
                   [qx.Command(qx.GateType.H, q) for q in range(n)])
    assert prog.kinds() == ["gates"] and all(isinstance(c, Command) for c in prog.kernels()[0][1])
    bound = [c.substitute(w.values) for c in w.commands]
    s_theirs = inspect_schedule([qx.from_ours(c) for c in bound], n, fusion_max_qubits=3, fusion_min_qubits=0)
    s_ours = inspect_schedule(bound, n, fusion_max_qubits=3, fusion_min_qubits=0)
    assert [(p, c.gate, c.qubits) for p, c in s_theirs] == [(p, c.gate, c.qubits) for p, c in s_ours]


def test_gate_given_by_name_or_number():
    assert as_command(qx.Command(qx.GateType.CX, 0, 1)).gate is GateType.CX

    class ByName:
        name = "CZ"
        value = "not a number"

    class Cmd:
        gate, qubits, params = ByName(), [0, 1], []

    assert as_command(Cmd()).gate is GateType.CZ
