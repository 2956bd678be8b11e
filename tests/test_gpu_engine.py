"""The engine layer on the GPU: primitives against the oracle, the three
gradient methods against each other and the oracle, the EXACT compaction
and the torch.autograd boundary."""
import numpy as np
import pytest

import _oracle as O
from openqarp_amd import workloads as W
from openqarp_amd._abi import pack_observable, pack_stream
from openqarp_amd.engine import HipEngine
from openqarp_amd.errors import CapabilityError
from openqarp_amd.ir import Command, GateType, Param
from openqarp_amd.grouping import FullyCommuting, NoGrouping, QubitWiseCommuting
from openqarp_amd.primitives import PauliAveraging, Sampler, StateVector, Target, diagonal_pauli_vector
from openqarp_amd.results import EXACT
from openqarp_amd.simulator import QarpSimulator
from test_host_api import _gradient_circuit

pytestmark = pytest.mark.gpu


def test_statevector_targets_match_oracle():
    n = 9
    rng = np.random.default_rng(1)
    ket = W.brickwork_haar(n, 4, 11)
    bra = W.brickwork_haar(n, 3, 12)
    obs = W.random_pauli_sum(n, 30, 7)
    eng = HipEngine()
    k = O.statevector(pack_stream(ket), n, 3, 12)
    b = O.statevector(pack_stream(bra), n, 3, 12)
    po = pack_observable(obs)
    ev = eng.run(StateVector(ket, n, obs))
    assert isinstance(ev, float) and abs(ev - O.transition(k, k, n, po).real) < 1e-12
    ov = eng.run(StateVector(ket, n, bra=bra, target=Target.OVERLAP))
    assert abs(ov - np.vdot(b, k)) < 1e-12
    tr = eng.run(StateVector(ket, n, obs, bra=bra, target=Target.TRANSITION_AMPLITUDE))
    assert abs(tr - O.transition(b, k, n, po)) < 1e-12


def test_z_only_expectation_matches_walsh_hadamard_diagonal():
    n = 12
    edges = W.random_regular_graph(3, n, 2)
    obs = [([(a, "Z"), (b, "Z")], 0.5) for a, b in edges] + [([], -0.5 * len(edges))]
    ket = W.brickwork_haar(n, 4, 3)
    k = O.statevector(pack_stream(ket), n, 3, 12)
    want = float(np.vdot(k, diagonal_pauli_vector(obs, n) * k).real)
    assert abs(HipEngine().run(StateVector(ket, n, obs)) - want) < 1e-12


def test_exact_sampler_matches_oracle_probabilities():
    n = 10
    circ = [Command(GateType.H, q) for q in range(0, n, 3)] + [Command(GateType.CX, q, q + 1) for q in range(n - 1)]
    circ += [Command(GateType.Ry, 4, 0.3)]
    sv = O.statevector(pack_stream(circ), n, 3, 12)
    p = np.abs(sv) ** 2
    (idx,) = np.nonzero(p > 1e-12)
    sim = QarpSimulator()
    raw = Sampler(circ, n, EXACT).raw(sim)
    assert np.array_equal(raw.keys, idx) and np.abs(raw.probs - p[idx]).max() < 1e-15
    d = Sampler(circ, n, EXACT, measured=[4, 0]).run(sim)
    want = {}
    for i in idx:
        key = (int((i >> 4) & 1), int(i & 1))
        want[key] = want.get(key, 0.0) + p[i]
    assert set(d) == set(want) and all(abs(d[k] - v) < 1e-14 for k, v in want.items())


def test_exact_compaction_at_22_qubits():
    n = 22
    sim = QarpSimulator()
    circ = W.brickwork_haar(n, 2, 4)
    st = sim.new_state(n)
    sim.apply(circ, st)
    probs = np.abs(st.download()) ** 2          # _exact_result's formula (_engine.py:100-102)
    for thr in (1e-12, 2e-7, 1e-6):
        keys, vals = sim.exact_state(st, thr)
        (idx,) = np.nonzero(probs > thr)
        assert np.array_equal(keys, idx) and np.array_equal(vals, probs[idx])
    # The sampler's probabilities are std::norm's x*x + y*y, within an ulp of these.
    assert np.abs(sim.probabilities_state(st) - probs).max() <= 2 * np.spacing(probs.max())


def test_counts_sampler_is_bit_exact_with_the_oracle():
    n, shots, seed = 9, 3000, 77
    circ = W.brickwork_haar(n, 3, 21)
    sv = O.statevector(pack_stream(circ), n, 3, 12)
    outcomes = O.sample(sv, n, shots, seed)
    d = Sampler(circ, n, shots, measured=[0, 3, 8]).run(QarpSimulator(), seed=seed)
    want = {}
    for o in outcomes.tolist():
        key = ((o >> 0) & 1, (o >> 3) & 1, (o >> 8) & 1)
        want[key] = want.get(key, 0) + 1
    assert {k: round(v * shots) for k, v in d.items()} == want


def test_pauli_averaging_exact_equals_expectation_and_sampling_is_close():
    n = 8
    ket = W.brickwork_haar(n, 3, 8)
    obs = W.random_pauli_sum(n, 25, 3) + [([], 0.4)]
    k = O.statevector(pack_stream(ket), n, 3, 12)
    want = O.transition(k, k, n, pack_observable(obs)).real
    for strategy in (FullyCommuting(), QubitWiseCommuting(), NoGrouping()):
        pa = PauliAveraging(ket, n, obs, EXACT, strategy)
        assert pa.n_groups == len(strategy.group(pa._terms, n))
        assert abs(HipEngine().run(pa) - want) < 1e-10
    est = HipEngine(seed=11).run(PauliAveraging(ket, n, obs, 20000))
    scale = sum(abs(complex(c)) for t, c in obs if t)
    assert abs(est - want) < 0.05 * scale


def test_pauli_averaging_shares_the_ket_state():
    """The ket runs once and each group samples a copy after its basis
    change: the estimate equals the one from running ket + basis change per
    group with the same seeds (the reference's per-group circuits)."""
    from openqarp_amd.simulator import QarpSimulator
    n = 9
    ket = W.brickwork_haar(n, 3, 4)
    obs = W.random_pauli_sum(n, 20, 6)
    sim = QarpSimulator()
    pa = PauliAveraging(ket, n, obs, 4000)
    seeds = list(range(100, 100 + pa.n_groups))
    want = pa.constant
    for g in range(pa.n_groups):
        want += pa.group_estimate(Sampler(pa.group_circuit(g), n, 4000).raw(sim, None, seeds[g]), g)
    assert pa.run(sim, None, seeds) == want


def test_gradient_methods_agree():
    n = 4
    cmds, obs, syms, params = _gradient_circuit(n)
    prim = StateVector(cmds, n, obs)
    eng = HipEngine()
    adj = eng.run_gradient(prim, params, syms, method="adjoint")
    ps = eng.run_gradient(prim, params, syms, method="parameter-shift")
    fd = eng.run_gradient(prim, params, syms, method="finite-diff")
    from openqarp_amd import gradients as G
    index = {s: i for i, s in enumerate(syms)}
    orc = O.run_gradient(pack_stream(G.rebase_differentiable(cmds), values=params, param_index=index), n,
                         pack_observable(obs), len(syms))
    assert np.abs(adj - orc).max() < 1e-10
    assert np.abs(ps - adj).max() < 1e-10
    assert np.abs(fd - adj).max() < 1e-8
    assert np.array_equal(eng.run_gradient(prim, params, syms), adj)
    with pytest.raises(ValueError):
        eng.run_gradient(prim, params, syms, method="spsa")


def test_parameter_shift_on_pauli_averaging_exact():
    n = 5
    w = W.hea(n, 2, 4, symbolic=True)
    obs = W.random_pauli_sum(n, 12, 1)
    params = dict(w.values)
    eng = HipEngine()
    adj = eng.run_gradient(StateVector(w.commands, n, obs), params, w.symbols)
    ps = eng.run_gradient(PauliAveraging(w.commands, n, obs, EXACT), params, w.symbols)
    assert np.abs(ps - adj).max() < 1e-9


def test_overlap_adjoint_gradient_matches_finite_differences():
    n = 5
    x, y = Param.symbol("x"), Param.symbol("y")
    ket = [Command(GateType.H, q) for q in range(n)] + [Command(GateType.Ry, 1, x), Command(GateType.CRz, 1, 2, y),
                                                        Command(GateType.RXX, 0, 3, x * 0.5)]
    bra = W.brickwork_haar(n, 2, 5) + [Command(GateType.Rz, 2, y)]
    prim = StateVector(ket, n, bra=bra, target=Target.OVERLAP)
    params = {"x": 0.7, "y": -0.4}
    eng = HipEngine()
    adj = eng.run_gradient(prim, params, ["x", "y"], method="adjoint")

    def f(p):
        return abs(eng.run(prim, p)) ** 2
    h = 1e-6
    fd = np.array([(f({**params, s: params[s] + h}) - f({**params, s: params[s] - h})) / (2 * h) for s in "xy"])
    assert np.abs(adj - fd).max() < 1e-8
    with pytest.raises(CapabilityError):
        eng.run_gradient(prim, params, ["x"], method="parameter-shift")


def test_autograd_function_backward_is_the_adjoint_gradient():
    torch = pytest.importorskip("torch")
    from openqarp_amd.autograd import expectation
    n = 6
    w = W.hea(n, 2, 9, symbolic=True)
    obs = W.tfim_chain(n, 0.7)
    sim = QarpSimulator()
    theta = torch.tensor([w.values[s] for s in w.symbols], dtype=torch.float64, requires_grad=True)
    e = expectation(theta, sim, w.commands, n, obs, w.symbols)
    e.backward()
    params = dict(w.values)
    want = np.asarray(sim.run_gradient(w.commands, n, obs, params, w.symbols))
    assert np.abs(theta.grad.numpy() - want).max() < 1e-12
    assert abs(e.item() - sim.expectation([c.substitute(params) for c in w.commands], n, obs)) < 1e-12
    with torch.no_grad():
        e2 = expectation(theta, sim, w.commands, n, obs, w.symbols)
    assert abs(e2.item() - e.item()) < 1e-12
    ok = torch.autograd.gradcheck(lambda t: expectation(t, sim, w.commands, n, obs, w.symbols),
                                  (theta.detach().clone().requires_grad_(True),), eps=1e-6, atol=1e-7)
    assert ok


def test_autograd_forward_does_no_gradient_work(monkeypatch):
    """The forward of the autograd function evaluates <H> only; the adjoint
    sweep runs when (and each time) backward is called."""
    torch = pytest.importorskip("torch")
    from openqarp_amd.autograd import expectation
    n = 5
    w = W.hea(n, 2, 4, symbolic=True)
    obs = W.tfim_chain(n, 0.9)
    sim = QarpSimulator()
    calls = []
    real = sim.run_gradient_energy
    monkeypatch.setattr(sim, "run_gradient_energy", lambda *a, **k: calls.append(1) or real(*a, **k))
    theta = torch.tensor([w.values[s] for s in w.symbols], dtype=torch.float64, requires_grad=True)
    e = expectation(theta, sim, w.commands, n, obs, w.symbols)
    assert calls == []
    (2.0 * e).backward()
    assert calls == [1]
    want = 2.0 * np.asarray(sim.run_gradient(w.commands, n, obs, dict(w.values), w.symbols))
    assert np.abs(theta.grad.numpy() - want).max() < 1e-12


def test_batch_expectation_against_oracle():
    """batch_expectation (qarp_simulator.cpp:971-984): one energy per
    parameter set, each equal to the oracle's statevector + transition."""
    n = 9
    w = W.c3_gradient(n, 2, 13, 25)
    sim = QarpSimulator()
    rng = np.random.default_rng(6)
    sets = [w.values] + [{s: float(rng.uniform(0, 6.3)) for s in w.symbols} for _ in range(4)]
    got = sim.batch_expectation(w.commands, n, w.observable, sets)
    po = pack_observable(w.observable)
    assert len(got) == len(sets)
    for e, vals in zip(got, sets):
        sv = O.statevector(pack_stream([c.substitute(vals) for c in w.commands]), n, 3, 12)
        assert abs(e - O.transition(sv, sv, n, po).real) < 1e-12
    assert sim.batch_expectation(w.commands, n, w.observable, []) == []


def test_pauli_averaging_bell_state_fully_commuting_single_group():
    """XX + YY + ZZ on the Bell state (H, CX): one FullyCommuting group whose
    Clifford reads all three terms at once (the reference's
    test_pauli_averaging.py:92-111): exactly 1 in the EXACT readout, and the
    sampled estimate equals the per-group circuit's bit for bit."""
    n = 2
    ket = [Command(GateType.H, 0), Command(GateType.CX, 0, 1)]
    obs = [([(0, "X"), (1, "X")], 1.0), ([(0, "Y"), (1, "Y")], 1.0), ([(0, "Z"), (1, "Z")], 1.0)]
    sim = QarpSimulator()
    pa = PauliAveraging(ket, n, obs, EXACT)
    assert pa.n_groups == 1
    assert abs(pa.run(sim) - 1.0) < 1e-12
    sampled = PauliAveraging(ket, n, obs, 3000)
    got = sampled.run(sim, None, [17])
    want = sampled.group_estimate(Sampler(sampled.group_circuit(0), n, 3000).raw(sim, None, 17), 0)
    assert got == want and abs(got - 1.0) < 1e-12   # the Bell state is an eigenstate: every shot agrees
