"""Host logic of the engine layer, on the CPU: result types, Pauli grouping,
the Z-only diagonal, and the parameter-shift / finite-difference stencils —
the stencils evaluated with oracle energies and compared with the oracle's
adjoint gradient, so a wrong shift, coefficient or affine scaling fails.
No GPU calls."""
import math

import numpy as np
import pytest

import _fixtures as F
import _oracle as O
from openqarp_amd import gradients as G
from openqarp_amd._abi import pack_observable, pack_stream
from openqarp_amd.errors import CapabilityError
from openqarp_amd.ir import Command, GateType, Param, SamplingResult
from openqarp_amd.primitives import decompose_observable, diagonal_pauli_vector
from openqarp_amd.results import EXACT, ExactResult, distribution_from_result, pack_bits


def test_pack_bits_and_marginal_distribution():
    out = np.array([0b101, 0b011, 0b110], dtype=np.int64)
    assert pack_bits(out, [0, 2]).tolist() == [0b11, 0b01, 0b10]
    res = SamplingResult(n_qubits=3, n_shots=4, counts={0b101: 2, 0b011: 1, 0b110: 1})
    d = distribution_from_result(res, [2, 0])    # key bit 0 <- qubit 2, bit 1 <- qubit 0
    assert d[(1, 1)] == 0.5 and d[(0, 1)] == 0.25 and d[(1, 0)] == 0.25
    assert sum(d.values()) == pytest.approx(1.0) and d.n_shots == 4
    er = ExactResult(3, np.array([1, 6]), np.array([0.25, 0.75]))
    de = distribution_from_result(er, [1, 2])
    assert de[(0, 0)] == 0.25 and de[(1, 1)] == 0.75 and de.n_shots is None
    assert EXACT.value == "exact"


def test_decompose_observable_folds_identity_and_keeps_real_parts():
    obs = [([(0, "X"), (1, "Z")], 1.0), ([(1, "Z")], 0.5), ([(0, "Y")], 2.0 + 0.5j), ([], 0.25),
           ([(2, "I")], 0.1), ([(0, "X"), (2, "Y")], -1.0)]
    const, terms, coeffs = decompose_observable(obs)
    assert const == pytest.approx(0.35)
    assert terms == [{0: "X", 1: "Z"}, {1: "Z"}, {0: "Y"}, {0: "X", 2: "Y"}]
    assert coeffs.tolist() == [1.0, 0.5, 2.0, -1.0]
    with pytest.raises(ValueError, match="repeated"):
        decompose_observable([([(0, "Z"), (0, "X")], 1.0)])
    with pytest.raises(ValueError, match="invalid Pauli"):
        decompose_observable([([(0, "Q")], 1.0)])


def test_diagonal_pauli_vector_is_the_operator_diagonal():
    n = 5
    rng = np.random.default_rng(3)
    obs = [([(int(a), "Z"), (int(b), "Z")], float(rng.normal())) for a, b in rng.integers(0, n, (7, 2))
           if a != b] + [([(2, "Z")], 0.3), ([], -1.5)]
    d = diagonal_pauli_vector(obs, n)
    M = F.pauli_matrix(obs, n)
    assert np.abs(np.diag(M) - d).max() < 1e-13 and np.abs(M - np.diag(np.diag(M))).max() == 0
    with pytest.raises(ValueError):
        diagonal_pauli_vector([([(0, "X")], 1.0)], n)


def test_shift_rules_recover_exact_derivatives_of_their_spectra():
    """Two-term: f = a + b cos(θ + φ) (frequency 1 in θ for a generator with
    spectrum ±½ acting as e^{-iθG} on both sides).  Four-term: frequencies ½
    and 1 (controlled rotations)."""
    rng = np.random.default_rng(0)
    for _ in range(5):
        a, b, c, p1, p2, th = rng.normal(size=6)
        f2 = lambda t: a + b * math.cos(t + p1)
        d2 = -b * math.sin(th + p1)
        assert sum(w * f2(th + s) for s, w in G.SHIFT_RULES[GateType.Rx]) == pytest.approx(d2, abs=1e-13)
        f4 = lambda t: a + b * math.cos(t / 2 + p1) + c * math.cos(t + p2)
        d4 = -0.5 * b * math.sin(th / 2 + p1) - c * math.sin(th + p2)
        assert sum(w * f4(th + s) for s, w in G.SHIFT_RULES[GateType.CRz]) == pytest.approx(d4, abs=1e-13)
    with pytest.raises(CapabilityError):
        G.shift_rule(GateType.H)


def _oracle_energy(commands, n, obs):
    po = pack_observable(obs)

    def evaluate(points):
        return [O.transition(s := O.statevector(pack_stream(commands, values=p), n, 0, 0), s, n, po).real
                for p in points]
    return evaluate


def _gradient_circuit(n=4):
    x, y, z = Param.symbol("x"), Param.symbol("y"), Param.symbol("z")
    cmds = [Command(GateType.H, q) for q in range(n)]
    cmds += [Command(GateType.Rx, 0, x * 2.0 + 0.1), Command(GateType.Ry, 1, y), Command(GateType.Rz, 2, z - 0.3),
             Command(GateType.RXX, 0, 1, x), Command(GateType.RYY, 1, 2, y * -0.5), Command(GateType.RZZ, 2, 3, z),
             Command(GateType.P, 3, x + y), Command(GateType.CP, 0, 3, z * 1.5), Command(GateType.CRx, 1, 0, x),
             Command(GateType.CRy, 2, 1, y * 3.0), Command(GateType.CRz, 3, 2, z + x),
             Command(GateType.U, 1, params=[x, y * 2.0, z])]
    cmds += [Command(GateType.CX, q, q + 1) for q in range(n - 1)]
    obs = [([(0, "Z"), (1, "X")], 0.7), ([(2, "Y")], -0.4), ([(3, "Z"), (0, "Y")], 0.9), ([(1, "Z")], 0.2)]
    return cmds, obs, ["x", "y", "z"], {"x": 0.37, "y": -1.1, "z": 2.2}


def test_parameter_shift_matches_adjoint_oracle():
    n = 4
    cmds, obs, syms, params = _gradient_circuit(n)
    plan = G.ShiftPlan(cmds, syms)
    ps = G.evaluate_stencils(_oracle_energy(plan.commands, n, obs), plan.base_point(params),
                             [plan.stencil(s) for s in syms])
    index = {s: i for i, s in enumerate(syms)}
    adj = O.run_gradient(pack_stream(G.rebase_differentiable(cmds), values=params, param_index=index), n,
                         pack_observable(obs), len(syms))
    assert np.abs(ps - adj).max() < 1e-10


def test_rebased_u_is_the_same_state():
    n = 2
    t, p, l = Param.symbol("t"), Param.symbol("p"), Param.symbol("l")
    cmds = [Command(GateType.H, 0), Command(GateType.U, 0, params=[t, p, l]), Command(GateType.CX, 0, 1)]
    vals = {"t": 0.4, "p": -1.3, "l": 2.1}
    a = O.statevector(pack_stream(cmds, values=vals), n, 0, 0)
    b = O.statevector(pack_stream(G.rebase_differentiable(cmds), values=vals), n, 0, 0)
    assert np.abs(a - b).max() < 1e-14
    with pytest.raises(CapabilityError):
        G.rebase_differentiable([Command(GateType.CU, 0, 1, params=[t, p, l, t])])


def test_finite_difference_stencils_and_occurrences():
    n = 4
    cmds, obs, syms, params = _gradient_circuit(n)
    fd = G.evaluate_stencils(_oracle_energy(cmds, n, obs), params, G.finite_difference_stencils(syms, 1e-5, 2))
    index = {s: i for i, s in enumerate(syms)}
    adj = O.run_gradient(pack_stream(G.rebase_differentiable(cmds), values=params, param_index=index), n,
                         pack_observable(obs), len(syms))
    assert np.abs(fd - adj).max() < 1e-8
    plan = G.ShiftPlan(cmds, syms)
    assert len(plan.occurrences["x"]) == 6 and {c for _, c, _ in plan.occurrences["x"]} == {2.0, 1.0}
    st = plan.stencil("x")
    assert ({plan.occurrences["x"][0][0]: math.pi / 4}, 1.0) in st   # Rx(2x + 0.1): shift π/4, weight 2·½
    with pytest.raises(ValueError):
        G.finite_difference_stencils(syms, 0.0)
