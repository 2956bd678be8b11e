"""Merged phase ops on the device: runs of four or more diagonal gates in a
pass (RZZ, CZ, CP, CRZ, CS, CSdg, two-qubit MCZ, RZ, P, S, Sdg, T, Tdg, Z,
GPhase, Custom diagonals) applied as one phase evaluation per amplitude, in
every single-state tile kernel (one tile per workgroup, the streaming kernel
of the widest tiles, batched instances), against the oracle: 1e-12 in
complex128, 2e-5 in complex64."""
import numpy as np
import pytest

import _oracle as O
from openqarp_amd import workloads as W
from openqarp_amd._abi import pack_observable, pack_stream
from openqarp_amd.simulator import CompiledCircuit, DeviceState, QarpSimulator
from test_schedule import phase_run_circuit

pytestmark = pytest.mark.gpu

TOL = {"complex128": 1e-12, "complex64": 2e-5}
# <C> and gradients (sums over the state): complex64 with FP64 reductions.
SUM_TOL = {"complex128": 1e-10, "complex64": 1e-3}


@pytest.mark.parametrize("precision", ["complex128", "complex64"])
@pytest.mark.parametrize("n,tile,low", [(5, 0, 0), (9, 0, 0), (12, 11, 3), (16, 12, 4), (18, 11, 4), (20, 12, 3)])
def test_phase_runs_match_oracle(n, tile, low, precision):
    cmds = phase_run_circuit(n, 4, 100 + n + tile)
    print(f"phase runs n={n} tile={tile} low={low} {precision}", flush=True)
    sim = QarpSimulator(precision=precision, tile_qubits=tile, low_qubits=low)
    sim.fusion_max_qubits, sim.fusion_min_qubits = 1, 0
    got = sim.statevector(cmds, n)
    ref = O.statevector(pack_stream(cmds), n, 0, 0)
    assert np.abs(got - ref).max() < TOL[precision]


@pytest.mark.parametrize("precision,tile", [("complex128", 13), ("complex64", 14), ("complex64", 13)])
def test_phase_runs_in_the_widest_tiles(precision, tile):
    """The widest tiles run on the streaming kernel (rows 8..13 of the tile
    carry terms)."""
    n = 18
    cmds = phase_run_circuit(n, 3, 7 + tile)
    print(f"widest tiles {precision} T={tile}", flush=True)
    sim = QarpSimulator(precision=precision, tile_qubits=tile)
    sim.fusion_max_qubits, sim.fusion_min_qubits = 1, 0
    got = sim.statevector(cmds, n)
    ref = O.statevector(pack_stream(cmds), n, 0, 0)
    assert np.abs(got - ref).max() < TOL[precision]


@pytest.mark.parametrize("precision", ["complex128", "complex64"])
def test_qaoa_at_fusion_one(precision):
    """QAOA MaxCut at 20 qubits with its cost layers merged: state, <C> and
    the adjoint gradient (whose forward sweep merges) against the oracle."""
    n = 20
    w = W.qaoa_maxcut(n, 4, 5)
    print(f"qaoa fusion 1 {precision}", flush=True)
    sim = QarpSimulator(precision=precision)
    sim.fusion_max_qubits, sim.fusion_min_qubits = 1, 0
    ref = O.statevector(pack_stream(w.commands), n, 0, 0)
    got = sim.statevector(w.commands, n)
    assert np.abs(got - ref).max() < TOL[precision]
    po = pack_observable(w.observable)
    c_ref = O.transition(ref, ref, n, po).real
    assert abs(sim.expectation(w.commands, n, w.observable) - c_ref) < SUM_TOL[precision]
    ws = W.qaoa_maxcut(n, 4, 5, symbolic=True)
    g, e = sim.run_gradient_energy(ws.commands, n, ws.observable, ws.values, ws.symbols)
    index = {s: i for i, s in enumerate(ws.symbols)}
    ps = pack_stream(ws.commands, values=ws.values, param_index=index)
    g_ref = O.run_gradient(ps, n, po, len(ws.symbols))
    assert np.abs(g - g_ref).max() < SUM_TOL[precision]
    assert abs(e - c_ref) < SUM_TOL[precision]


def test_compiled_qaoa_rebinds_merged_phases():
    """A compiled QAOA circuit at fusion 1: rebinding gamma / beta rewrites
    the merged ops' factors in place (no rebuild), graph replays included."""
    n = 16
    ws = W.qaoa_maxcut(n, 3, 9, symbolic=True)
    sim = QarpSimulator()
    cc = CompiledCircuit(sim, ws.commands, n, fusion_max_qubits=1, values=ws.values)
    st = DeviceState(n, "complex128")
    for trial in range(3):
        vals = {k: v + 0.37 * trial for k, v in ws.values.items()}
        if trial:
            assert cc.update(vals) is False
        st.reset()
        (cc.apply_graph if trial == 2 else cc.apply)(st)
        ref = O.statevector(pack_stream(ws.commands, values=vals), n, 0, 0)
        assert np.abs(st.download() - ref).max() < 1e-12, trial


def test_batched_instances_with_phase_runs():
    """Batched forward steps (tile_batch_kernel) carry merged phase ops:
    energies and gradients equal each instance's single run."""
    sim = QarpSimulator()
    sim.fusion_max_qubits, sim.fusion_min_qubits = 1, 0
    inst = []
    for i, n in enumerate((9, 12, 14, 16)):
        ws = W.qaoa_maxcut(n, 2 + i % 2, 20 + i, symbolic=True)
        inst.append((ws.commands, n, ws.observable, ws.values, ws.symbols))
    energies, grads = sim.batch_energy_grad(inst)
    for i, (cmds, n, obs, params, order) in enumerate(inst):
        g, e = sim.run_gradient_energy(cmds, n, obs, params, order)
        assert abs(energies[i] - e) < 1e-12
        assert np.abs(grads[i] - g).max() < 1e-10
        index = {s: k for k, s in enumerate(order)}
        ps = pack_stream(cmds, values=params, param_index=index)
        sv = O.statevector(ps, n, 0, 0)
        assert abs(e - O.transition(sv, sv, n, pack_observable(obs)).real) < 1e-12


def diagonal_gradient_circuit(n, layers, seed):
    """Mixer layers around runs of parametric diagonal gates sharing a few
    symbols (affine angles, some with two symbols), with unparametrised CZ,
    S and T gates in the runs: the reverse sweep merges such runs."""
    from openqarp_amd.ir import Command, GateType, Param
    rng = np.random.default_rng(seed)
    syms = [Param.symbol(f"t{i}") for i in range(2 * layers + 1)]
    values = {f"t{i}": float(v) for i, v in enumerate(rng.uniform(-2, 2, len(syms)))}
    cmds = [Command(GateType.H, q) for q in range(n)]
    for layer in range(layers):
        a, b = syms[2 * layer], syms[2 * layer + 1]
        for _ in range(3 * n):
            r = rng.random()
            q0, q1 = (int(x) for x in rng.choice(n, 2, replace=False))
            w = float(rng.uniform(0.3, 1.7))
            if r < 0.35:
                cmds.append(Command(GateType.RZZ, q0, q1, a * w))
            elif r < 0.5:
                cmds.append(Command(GateType.CRz, q0, q1, a * w + b * 0.5))
            elif r < 0.6:
                cmds.append(Command(GateType.CP, q0, q1, b * w))
            elif r < 0.75:
                cmds.append(Command(GateType.Rz, q0, a * w - 0.2))
            elif r < 0.85:
                cmds.append(Command(GateType.P, q0, b * w))
            elif r < 0.9:
                cmds.append(Command(GateType.GPhase, params=[a * 0.7]))
            elif r < 0.95:
                cmds.append(Command(GateType.CZ, q0, q1))
            else:
                cmds.append(Command([GateType.S, GateType.T][int(rng.integers(2))], q0))
        for q in range(n):
            cmds.append(Command(GateType.Rx, q, syms[-1] * float(rng.uniform(0.5, 1.5))))
    names = [f"t{i}" for i in range(len(syms))]
    obs = W.random_pauli_sum(n, 12, seed + 1)
    return cmds, values, names, obs


@pytest.mark.parametrize("precision", ["complex128", "complex64"])
@pytest.mark.parametrize("n,tile", [(6, 0), (12, 0), (17, 0), (17, 12)])
def test_adjoint_merges_diagonal_runs(n, tile, precision):
    """Runs of parametric diagonal gates in the reverse sweep: one inner
    product per symbol before the run's merged un-applies; the gradient
    equals the oracle's (1e-10 in complex128)."""
    cmds, values, names, obs = diagonal_gradient_circuit(n, 3, 40 + n)
    print(f"adjoint diagonal runs n={n} tile={tile} {precision}", flush=True)
    sim = QarpSimulator(precision=precision, tile_qubits=tile)
    g, e = sim.run_gradient_energy(cmds, n, obs, values, names)
    index = {s: i for i, s in enumerate(names)}
    ps = pack_stream(cmds, values=values, param_index=index)
    g_ref = O.run_gradient(ps, n, pack_observable(obs), len(names))
    tol = 1e-10 if precision == "complex128" else 1e-3
    assert np.abs(g - g_ref).max() < tol
    sv = O.statevector(ps, n, 0, 0)
    assert abs(e - O.transition(sv, sv, n, pack_observable(obs)).real) < tol


def test_adjoint_phi_with_merged_runs():
    """run_gradient_phi shares the reverse sweep: a caller-supplied phi
    against the oracle's run_gradient_phi."""
    n = 11
    cmds, values, names, _ = diagonal_gradient_circuit(n, 2, 9)
    rng = np.random.default_rng(3)
    phi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    sim = QarpSimulator()
    g = np.array(sim.run_gradient_phi(cmds, n, phi, values, names))
    index = {s: i for i, s in enumerate(names)}
    ps = pack_stream(cmds, values=values, param_index=index)
    g_ref = O.run_gradient_phi(ps, n, phi, len(names))
    assert np.abs(g - g_ref).max() < 1e-10 * max(1.0, np.abs(g_ref).max())


@pytest.mark.parametrize("precision", ["complex128", "complex64"])
@pytest.mark.parametrize("fusion,tile,low", [(0, 6, 2), (1, 9, 4), (2, 11, 3), (3, 12, 4), (4, 12, 3), (5, 10, 1)])
def test_phase_runs_inside_random_circuits(fusion, tile, low, precision):
    """Phase runs between random gates of every kind (dense, Custom, MCZ,
    markers) at several fusion widths and tile shapes, with qubit relabeling
    on: the merged ops address tile positions, row bits and relabeled global
    bits as the pass's layout says."""
    from test_gpu_parity import random_circuit
    n = 17
    rng = np.random.default_rng(fusion * 31 + tile)
    cmds = []
    for part in range(4):
        cmds += random_circuit(n, 25, int(rng.integers(1 << 30)))
        cmds += [c for c in phase_run_circuit(n, 1, int(rng.integers(1 << 30)))[n:]
                 if c.gate.name not in ("Rx", "Ry")]
    print(f"random + phase runs fusion={fusion} tile={tile} low={low} {precision}", flush=True)
    sim = QarpSimulator(precision=precision, tile_qubits=tile, low_qubits=low)
    sim.fusion_max_qubits, sim.fusion_min_qubits = fusion, 0
    got = sim.statevector(cmds, n)
    ref = O.statevector(pack_stream(cmds), n, fusion, 0)
    assert np.abs(got - ref).max() < TOL[precision]
