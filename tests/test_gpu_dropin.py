"""The simulator called the way QarpEngine calls qx.QarpSimulator
(_qarp_engine.py:398-433, state_vector.py:380-403, _gradients.py:519-544),
with streams of objects shaped like qx.Command / qx.Param (tests/_qx_stub.py):
every method returns what it returns for this package's own Commands, bit
for bit, and the values agree with the oracle."""
import numpy as np
import pytest

import _oracle as O
import _qx_stub as qx
from openqarp_amd import workloads as W
from openqarp_amd._abi import pack_observable, pack_stream
from openqarp_amd.ir import as_commands
from openqarp_amd.program import Program
from openqarp_amd.simulator import QarpSimulator

pytestmark = pytest.mark.gpu


def test_engine_calls_with_reference_shaped_commands():
    n = 10
    w = W.c3_gradient(n, 3, 21, 30)
    theirs = [qx.from_ours(c) for c in w.commands]
    sim = QarpSimulator()
    bound_ours = [c.substitute(w.values) for c in w.commands]
    bound_theirs = [qx.from_ours(c) for c in bound_ours]
    sv = sim.statevector(bound_theirs, n)
    assert np.array_equal(sv, sim.statevector(bound_ours, n))
    ref = O.statevector(pack_stream(bound_ours), n, 3, 12)
    assert np.abs(sv - ref).max() < 1e-12
    assert sim.expectation(bound_theirs, n, w.observable) == sim.expectation(bound_ours, n, w.observable)
    g1 = sim.run_gradient(theirs, n, w.observable, w.values, w.symbols)
    g2 = sim.run_gradient(w.commands, n, w.observable, w.values, w.symbols)
    assert g1 == g2
    phi = np.vdot(ref, sv) * ref
    assert sim.run_gradient_phi(theirs, n, phi, w.values, w.symbols) == \
        sim.run_gradient_phi(w.commands, n, phi, w.values, w.symbols)
    sets = [w.values, {s: v + 0.1 for s, v in w.values.items()}]
    assert sim.batch_expectation(theirs, n, w.observable, sets) == \
        sim.batch_expectation(w.commands, n, w.observable, sets)
    measured = bound_theirs + [qx.Command(qx.GateType.Measure, q, cbits=[q]) for q in range(n)]
    r1 = sim.run(measured, n, 500, 17)
    r2 = sim.run(as_commands(measured), n, 500, 17)
    assert r1.counts == r2.counts and r1.cbit_history == r2.cbit_history
    assert sum(r1.counts.values()) == 500


def test_program_run_with_reference_shaped_gates():
    n = 9
    cmds = W.brickwork_haar(n, 2, 5)
    rng = np.random.default_rng(2)
    layer = [qx.Command(qx.GateType.Ry, q, float(rng.uniform(0, 6))) for q in range(n)]
    p1, p2 = Program(), Program()
    p1.add_gates(cmds)
    p1.add_gates(layer)
    p2.add_gates(cmds)
    p2.add_gates(as_commands(layer))
    sim = QarpSimulator()
    assert np.array_equal(sim.program_statevector(p1, n), sim.program_statevector(p2, n))
    assert sim.program_run(p1, n, 300, 4).counts == sim.program_run(p2, n, 300, 4).counts
    sv = sim.program_statevector(p1, n)
    po = pack_observable([([(0, "Z")], 1.0)])
    assert abs(sim.expectation_state(sim.new_state(n, sv), [([(0, "Z")], 1.0)]) -
               O.transition(sv, sv, n, po).real) < 1e-12
