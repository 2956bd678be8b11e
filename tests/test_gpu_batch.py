"""Batched instances (qhip_batch_energy_grad): every instance's energy and
adjoint gradient must equal its own single run (complex128, 1e-12 / 1e-10)
and the oracle; complex64 states with FP64 reductions stay within float32
accuracy of complex128."""
import numpy as np
import pytest

import _oracle as O
from openqarp_amd import workloads as W
from openqarp_amd._abi import pack_observable, pack_stream
from openqarp_amd.simulator import QarpSimulator

pytestmark = pytest.mark.gpu


def _instances(m, seed, n_min=5, n_max=13, d_min=1, d_max=5):
    ws = W.batched_vqe(m, seed, n_min, n_max, d_min, d_max)
    return [(w.commands, w.n_qubits, w.observable, w.values, w.symbols) for w in ws]


def test_batch_matches_single_runs_and_oracle():
    inst = _instances(24, 3)
    sim = QarpSimulator()
    energies, grads = sim.batch_energy_grad(inst)
    for i, (cmds, n, obs, params, order) in enumerate(inst):
        g, e = sim.run_gradient_energy(cmds, n, obs, params, order)
        assert abs(energies[i] - e) < 1e-12
        assert np.abs(grads[i] - g).max() < 1e-10
    for i in (0, 7, 23):
        cmds, n, obs, params, order = inst[i]
        index = {s: k for k, s in enumerate(order)}
        ps = pack_stream(cmds, values=params, param_index=index)
        sv = O.statevector(ps, n, 3, 12)
        po = pack_observable(obs)
        assert abs(energies[i] - O.transition(sv, sv, n, po).real) < 1e-12
        assert np.abs(grads[i] - O.run_gradient(ps, n, po, len(order))).max() < 1e-10


def test_batch_energies_only_and_unbatched_instances():
    inst = _instances(10, 8)
    # A 9-qubit brickwork fused at width 7 has a pass the batched kernel does
    # not run (dense op wider than five qubits): it takes the one-by-one path.
    wide = W.brickwork_haar(9, 6, 2)
    inst.append((wide, 9, W.random_pauli_sum(9, 10, 4), {}, []))
    sim = QarpSimulator()
    sim.fusion_max_qubits, sim.fusion_min_qubits = 7, 0
    energies, none = sim.batch_energy_grad(inst, gradients=False)
    assert none is None
    for i, (cmds, n, obs, params, order) in enumerate(inst):
        want = sim.expectation([c.substitute(params) if c.is_parametric() else c for c in cmds], n, obs)
        assert abs(energies[i] - want) < 1e-12
    e2, g2 = sim.batch_energy_grad(inst)
    assert np.abs(e2 - energies).max() < 1e-12 and g2[-1].size == 0


def test_complex64_batch_error_is_float32_sized():
    inst = _instances(16, 5, 8, 14, 2, 6)
    e128, g128 = QarpSimulator().batch_energy_grad(inst)
    e64, g64 = QarpSimulator(precision="complex64").batch_energy_grad(inst)
    for i in range(len(inst)):
        scale = sum(abs(complex(c)) for _, c in inst[i][2])
        assert abs(e64[i] - e128[i]) < 2e-5 * scale
        assert np.abs(g64[i] - g128[i]).max() < 2e-5 * scale


def test_qaoa_complex64_against_oracle():
    """The QAOA MaxCut workload in complex64: <C> and the adjoint gradient
    against the complex128 oracle (float32 accuracy), and shots drawn from
    the complex64 state bit-exact with the oracle sampler on the same
    amplitudes."""
    n = 16
    w = W.qaoa_maxcut(n, 3, 5, symbolic=True)
    sim = QarpSimulator(precision="complex64")
    index = {s: k for k, s in enumerate(w.symbols)}
    ps = pack_stream(w.commands, values=w.values, param_index=index)
    po = pack_observable(w.observable)
    sv = O.statevector(ps, n, 3, 12)
    e_ref = O.transition(sv, sv, n, po).real
    g_ref = O.run_gradient(ps, n, po, len(w.symbols))
    g, e = sim.run_gradient_energy(w.commands, n, w.observable, w.values, w.symbols)
    assert abs(e - e_ref) < 1e-4 and np.abs(g - g_ref).max() < 1e-4
    st = sim.new_state(n)
    sim.apply([c.substitute(w.values) for c in w.commands], st)
    amps = st.download()
    assert abs(sim.expectation_state(st, w.observable) - O.transition(amps, amps, n, po).real) < 1e-9
    assert np.array_equal(sim.sample_state(st, 5000, 123), O.sample(amps, n, 5000, 123))


def test_batch_session_graph_sweeps_match_direct_runs():
    """A session replayed as a hipGraph over parameter updates: every replay
    equals a fresh direct batch at the same values, with other device work
    between replays; a structure change rebuilds the session."""
    from openqarp_amd.simulator import BatchSession
    inst = _instances(12, 21)
    sim = QarpSimulator()
    sess = BatchSession(sim, inst)
    rng = np.random.default_rng(3)
    for sweep in range(3):
        if sweep:
            vals = [{s: float(rng.uniform(0, 6.28)) for s in order} for *_, order in inst]
            assert sess.update(vals) is False
            inst = [(c, n, o, v, order) for (c, n, o, _, order), v in zip(inst, vals)]
        for use_graph in (True, True, False):
            e, g = sess.run(use_graph=use_graph)
            e_ref, g_ref = sim.batch_energy_grad(inst)
            assert np.array_equal(e, e_ref)
            assert all(np.array_equal(a, b) for a, b in zip(g, g_ref))
            sim.expectation(W.brickwork_haar(8, 2, 1), 8, W.random_pauli_sum(8, 5, 2))   # other work
    zeros = [{s: 0.0 for s in order} for *_, order in inst]
    assert sess.update(zeros) is True
    e, g = sess.run()
    inst0 = [(c, n, o, v, order) for (c, n, o, _, order), v in zip(inst, zeros)]
    e_ref, g_ref = sim.batch_energy_grad(inst0)
    assert np.abs(e - e_ref).max() < 1e-12 and all(np.abs(a - b).max() < 1e-12 for a, b in zip(g, g_ref))
