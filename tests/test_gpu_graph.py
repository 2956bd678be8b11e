"""Compiled plans replayed as hipGraphs and rebound to new parameter values.
Every replay is checked against the direct launches (same kernels: bit-equal)
with other device work between replays, and rebinding against the oracle."""
import numpy as np
import pytest

import _oracle as O
from openqarp_amd import workloads as W
from openqarp_amd._abi import pack_observable, pack_stream
from openqarp_amd.simulator import CompiledCircuit, DeviceState, QarpSimulator

pytestmark = pytest.mark.gpu


def test_graph_replays_equal_direct_launches():
    n = 14
    cmds = W.brickwork_haar(n, 10, 6)
    sim = QarpSimulator(tile_qubits=6, low_qubits=2)
    cc = CompiledCircuit(sim, cmds, n, fusion_max_qubits=2)
    st = DeviceState(n)
    cc.apply(st)
    eager = st.download()
    assert cc.stats["passes"] > 5
    other = DeviceState(n)
    obs = W.random_pauli_sum(n, 20, 1)
    for rep in range(4):
        st.reset()
        cc.apply_graph(st)
        assert np.array_equal(st.download(), eager), f"replay {rep}"
        other.reset()
        sim.apply(cmds, other)                       # eager work between replays
        sim.expectation_state(other, obs)
    # A second state re-captures and gives the same amplitudes.
    other.reset()
    cc.apply_graph(other)
    assert np.array_equal(other.download(), eager)
    ref = O.statevector(pack_stream(cmds), n, 2, 0)
    assert np.abs(eager - ref).max() < 1e-12


@pytest.mark.parametrize("precision", ["complex128", "complex64"])
def test_rebinding_parameters_keeps_or_rebuilds_the_graph(precision):
    n = 12
    w = W.hea(n, 4, 13, symbolic=True)
    sim = QarpSimulator(precision=precision)
    tol = 1e-12 if precision == "complex128" else 2e-6
    cc = CompiledCircuit(sim, w.commands, n, fusion_max_qubits=3, values=w.values)
    st = DeviceState(n, precision)
    index = {s: k for k, s in enumerate(w.symbols)}
    rng = np.random.default_rng(0)
    for trial in range(3):
        vals = w.values if trial == 0 else {s: float(rng.uniform(0, 6.28)) for s in w.symbols}
        if trial:
            assert cc.update(vals) is False          # same structure: coefficients only
        for _ in range(2):
            st.reset()
            cc.apply_graph(st)
            ref = O.statevector(pack_stream(w.commands, values=vals, param_index=index), n, 3, 0)
            assert np.abs(st.download() - ref).max() < tol
    # Rotations only: at angle 0 every fused block is the identity and is
    # dropped, so the schedule changes and the tables are rebuilt.
    rot = [c for c in w.commands if c.is_parametric()]
    cr = CompiledCircuit(sim, rot, n, fusion_max_qubits=3, values=w.values)
    zeros = {s: 0.0 for s in w.symbols}
    assert cr.update(zeros) is True
    st.reset()
    cr.apply_graph(st)
    ref = O.statevector(pack_stream(rot, values=zeros, param_index=index), n, 3, 0)
    assert np.abs(st.download() - ref).max() < tol
    assert cr.update(w.values) is True
    st.reset()
    cr.apply_graph(st)
    ref = O.statevector(pack_stream(rot, values=w.values, param_index=index), n, 3, 0)
    assert np.abs(st.download() - ref).max() < tol
