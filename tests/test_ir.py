"""Host-side IR: constructors, dagger rules (core/command.cpp:19-59), affine
parameters (core/param.cpp:462-497) and the stream packing.  CPU only."""
import math

import numpy as np
import pytest

from openqarp_amd import workloads as W
from openqarp_amd._abi import OP_DTYPE, pack_observable, pack_stream
from openqarp_amd.errors import CapabilityError
from openqarp_amd.ir import Command, GateType, Param, affine_coefficients, cbit_register_width


def test_positional_constructor_forms():
    assert Command(GateType.X, 3).qubits == (3,)
    assert Command(GateType.CX, 0, 1).qubits == (0, 1)
    c = Command(GateType.Rx, 2, 0.5)
    assert c.qubits == (2,) and c.params[0].value() == 0.5
    c = Command(GateType.RZZ, 0, 4, Param.symbol("t"))
    assert c.qubits == (0, 4) and c.is_parametric()
    c = Command(GateType.U, 0, params=[Param.symbol("a"), 0.1, 0.2])
    assert c.qubits == (0,) and len(c.params) == 3


@pytest.mark.parametrize("g,inv", [(GateType.S, GateType.Sdg), (GateType.T, GateType.Tdg),
                                   (GateType.SX, GateType.SXdg), (GateType.iSWAP, GateType.iSWAPdg),
                                   (GateType.CS, GateType.CSdg), (GateType.CSX, GateType.CSXdg)])
def test_named_inverse_pairs(g, inv):
    nq = 2 if g.value >= 16 else 1
    c = Command(g, qubits=list(range(nq)))
    assert c.dagger().gate == inv and c.dagger().dagger().gate == g


def test_parametric_daggers():
    assert Command(GateType.Rx, 0, 0.3).dagger().params[0].value() == -0.3
    u = Command(GateType.U, 0, params=[0.1, 0.2, 0.3]).dagger()
    assert [p.value() for p in u.params] == [-0.1, -0.3, -0.2]
    cu = Command(GateType.CU, 0, 1, params=[0.1, 0.2, 0.3, 0.4]).dagger()
    assert [p.value() for p in cu.params] == [-0.1, -0.3, -0.2, -0.4]
    m = W.haar_unitary(4, np.random.default_rng(0))
    c = Command(GateType.Custom, qubits=[0, 1], unitary=m).dagger()
    assert np.allclose(c.unitary, m.conj().T)
    for g in (GateType.X, GateType.H, GateType.CX, GateType.ECR, GateType.CCX):
        nq = 3 if g == GateType.CCX else 2 if g.value >= 16 else 1
        assert Command(g, qubits=list(range(nq))).dagger().gate == g


def test_affine_parameters():
    a, b = Param.symbol("a"), Param.symbol("b")
    p = (a + b) * 0.5 - 0.25
    assert affine_coefficients(p) == {"a": 0.5, "b": 0.5}
    assert p.evaluate({"a": 1.0, "b": 3.0}) == pytest.approx(1.75)
    assert p.substitute({"a": 1.0}).free_symbols() == ["b"]
    assert affine_coefficients(2.0) == {}
    sq = Param.compound(lambda v: v["a"] ** 2, ["a"])
    with pytest.raises(CapabilityError):
        affine_coefficients(sq)
    lin = Param.compound(lambda v: 3 * v["a"] - v["b"], ["a", "b"])
    assert affine_coefficients(lin) == {"a": pytest.approx(3.0), "b": pytest.approx(-1.0)}


def test_pack_stream_layout_and_deps():
    cmds = [Command(GateType.Ry, 1, Param.symbol("x") * 2.0 + 0.1), Command(GateType.CX, 0, 1),
            Command(GateType.Custom, qubits=[1], unitary=np.eye(2)),
            Command(GateType.Measure, 0, cbits=[3])]
    ps = pack_stream(cmds, {"x": 0.4}, {"x": 0})
    assert ps.ops.dtype == OP_DTYPE and ps.ops.itemsize == 80
    assert ps.ops[0]["params"][0] == pytest.approx(0.9)
    assert ps.ops[0]["n_deps"] == 1 and ps.dep_index[0] == 0 and ps.dep_alpha[0] == 2.0
    assert list(ps.qubits) == [1, 0, 1, 1, 0]
    assert ps.ops[2]["unitary_offset"] == 0 and ps.unitaries.size == 8
    assert ps.ops[3]["cbit"] == 3 and ps.ops[1]["cbit"] == -1
    assert cbit_register_width(cmds) == 4
    with pytest.raises(RuntimeError, match="unresolved symbolic"):
        pack_stream(cmds)


def test_pack_observable():
    po = pack_observable([([(0, "X"), (3, "Y")], 0.5 + 0.25j), ([], 1.0)])
    assert po.terms[0]["n_letters"] == 2 and po.terms[1]["n_letters"] == 0
    assert bytes(po.letters) == b"XY"
    assert po.terms[0]["coeff_im"] == 0.25


def test_workloads_are_deterministic_and_shaped():
    a, b = W.brickwork_haar(30, 30, 2), W.brickwork_haar(30, 30, 2)
    assert len(a) == 450 and all(np.array_equal(x.unitary, y.unitary) for x, y in zip(a, b))
    for layer in range(30):
        qs = sorted(q for c in a[15 * layer:15 * layer + 15] for q in c.qubits)
        assert qs == list(range(30))
    u = a[0].unitary
    assert np.allclose(u @ u.conj().T, np.eye(4))
    c3 = W.c3_gradient(24, 8, 3, 1000)
    assert len(c3.symbols) == 384 and len(c3.observable) == 1000
    assert sum(c.gate == GateType.CZ for c in c3.commands) == 8 * 23
    edges = W.random_regular_graph(3, 34, 5)
    deg = np.bincount(np.array(edges).ravel(), minlength=34)
    assert len(edges) == 51 and (deg == 3).all()
