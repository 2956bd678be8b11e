"""The device scheduler and lowering, checked on the host: the scheduled
operation sequence (every pass, in execution order, as global operations)
must produce the same state as the input stream — evaluated with the oracle.
Exercises tiny tiles, scattered tile qubits, diagonal ops that skip the tile
constraint, wide blocks and both precisions.  No GPU calls."""
import numpy as np
import pytest

import _oracle as O
from openqarp_amd import workloads as W
from openqarp_amd._abi import pack_stream
from openqarp_amd.simulator import inspect_schedule
from test_gpu_parity import random_circuit


def scheduled_state(cmds, n, **kw):
    sched = inspect_schedule(cmds, n, **kw)
    passes = [p for p, _ in sched]
    assert passes == sorted(passes)
    return O.statevector(pack_stream([c for _, c in sched]), n, 0, 0), sched


@pytest.mark.parametrize("n", [2, 4, 7, 10, 12])
@pytest.mark.parametrize("fusion", [0, 2, 3, 5])
@pytest.mark.parametrize("tile,low", [(0, 0), (5, 2), (6, 4), (8, 1), (9, 6)])
def test_schedule_preserves_the_circuit(n, fusion, tile, low):
    cmds = random_circuit(n, 70, 31 * n + fusion) + W.brickwork_haar(n, 3, n)
    ref = O.statevector(pack_stream(cmds), n, fusion, 0)
    got, sched = scheduled_state(cmds, n, fusion_max_qubits=fusion, fusion_min_qubits=0, tile_qubits=tile,
                                 low_qubits=low)
    assert np.abs(got - ref).max() < 1e-12


def test_wide_blocks_get_their_own_passes():
    n = 10
    cmds = W.brickwork_haar(n, 6, 4)
    ref = O.statevector(pack_stream(cmds), n, 7, 0)
    got, sched = scheduled_state(cmds, n, fusion_max_qubits=7, fusion_min_qubits=0)
    assert np.abs(got - ref).max() < 1e-12
    assert any(len(c.qubits) > 5 for _, c in sched)


def test_diagonal_ops_need_no_tile_qubits():
    """A CZ / RZZ / MCZ ladder over all qubits fits in one pass of any tile."""
    from openqarp_amd.ir import Command, GateType
    n = 12
    cmds = [Command(GateType.H, q) for q in range(n)]
    diag = [Command(GateType.CZ, q, (q + 5) % n) for q in range(n)]
    diag += [Command(GateType.RZZ, q, (q + 7) % n, 0.3 * q) for q in range(n)]
    diag += [Command(GateType.MCZ, qubits=[0, 5, 9, 11])]
    _, sched = scheduled_state(cmds, n, fusion_max_qubits=0, tile_qubits=6)
    _, sched2 = scheduled_state(cmds + diag, n, fusion_max_qubits=0, tile_qubits=6)
    assert max(p for p, _ in sched2) == max(p for p, _ in sched)
    ref = O.statevector(pack_stream(cmds + diag), n, 0, 0)
    got, _ = scheduled_state(cmds + diag, n, fusion_max_qubits=0, tile_qubits=6)
    assert np.abs(got - ref).max() < 1e-12


def test_complex64_schedule_matches():
    n = 11
    cmds = random_circuit(n, 60, 5)
    ref = O.statevector(pack_stream(cmds), n, 3, 0)
    got, _ = scheduled_state(cmds, n, fusion_max_qubits=3, fusion_min_qubits=0, precision="complex64")
    assert np.abs(got - ref).max() < 1e-12


def test_thirty_qubit_brickwork_schedule_shape():
    """The benchmark circuit's schedule: every gate lands in some pass and the
    pass count stays far below the fused-op count."""
    cmds = W.brickwork_haar(30, 30, 2)
    for k in (2, 5):
        sched = inspect_schedule(cmds, 30, fusion_max_qubits=k, fusion_min_qubits=12)
        passes = max(p for p, _ in sched) + 1
        assert passes < len(sched)


def test_long_circuit_beyond_the_candidate_window():
    """More ops than the scheduler's candidate window (512): the passes still
    reproduce the circuit."""
    n = 9
    cmds = random_circuit(n, 900, 77) + W.brickwork_haar(n, 40, 5)
    ref = O.statevector(pack_stream(cmds), n, 2, 0)
    got, sched = scheduled_state(cmds, n, fusion_max_qubits=2, fusion_min_qubits=0, tile_qubits=5, low_qubits=2)
    assert len(sched) > 512
    assert np.abs(got - ref).max() < 1e-12


def test_frontier_scheduling_pass_counts():
    """Best-fit pass selection over the dependency frontier on the 30-qubit
    brickwork: at fusion width 2 and 2^12-amplitude tiles with 3 contiguous
    low qubits it needs at most 45 passes for the 436 fused ops (taking ops
    strictly in program order needed 62)."""
    cmds = W.brickwork_haar(30, 30, 2)
    sched = inspect_schedule(cmds, 30, fusion_max_qubits=2, fusion_min_qubits=0, tile_qubits=12, low_qubits=3)
    assert max(p for p, _ in sched) + 1 <= 45


def _swaps(sched):
    swap = np.eye(4)[[0, 2, 1, 3]]
    return sum(1 for _, c in sched if c.unitary is not None and c.unitary.shape == (4, 4)
               and np.array_equal(c.unitary, swap))


@pytest.mark.parametrize("n,tile,low,fusion", [(12, 7, 3, 3), (12, 8, 3, 2), (14, 9, 3, 4), (10, 6, 2, 2),
                                               (12, 6, 4, 2)])
def test_relabeling_passes_reproduce_the_circuit(n, tile, low, fusion):
    """Passes whose stores move the next pass's qubits into the contiguous
    low tile positions (reported as SWAPs of the pass's global bits after its
    ops): replayed in order they give the circuit's state, so the layout is
    back to the identity when the schedule ends.  Without relabeling no pass
    permutes."""
    cmds = W.brickwork_haar(n, 8, 3) + random_circuit(n, 40, 9)
    ref = O.statevector(pack_stream(cmds), n, fusion, 0)
    kw = dict(fusion_max_qubits=fusion, fusion_min_qubits=0, tile_qubits=tile, low_qubits=low)
    got, sched = scheduled_state(cmds, n, **kw)
    assert _swaps(sched) > 0
    assert np.abs(got - ref).max() < 1e-12
    got0, sched0 = scheduled_state(cmds, n, relabel_qubits=0, **kw)
    assert np.abs(got0 - ref).max() < 1e-12


def test_relabeling_saves_passes_on_the_brickwork():
    """30-qubit brickwork, fusion width 4, 3 contiguous low qubits: with
    relabeling 2^11-amplitude tiles need 56 passes instead of 62."""
    cmds = W.brickwork_haar(30, 30, 2)
    kw = dict(fusion_max_qubits=4, fusion_min_qubits=0, tile_qubits=11, low_qubits=3)
    on = inspect_schedule(cmds, 30, **kw)
    off = inspect_schedule(cmds, 30, relabel_qubits=0, **kw)
    assert max(p for p, _ in on) + 1 <= 56
    assert max(p for p, _ in off) + 1 == 62
    assert _swaps(off) == 0


def phase_run_circuit(n, layers, seed, run_len=None):
    """Layers of H / RX / RY mixers around runs of diagonal gates of every
    kind that merges (RZZ, CZ, CP, CRZ, CS, CSdg, 2-qubit MCZ, RZ, P, S, Sdg,
    T, Tdg, Z, GPhase, a Custom diagonal unitary)."""
    from openqarp_amd.ir import Command, GateType
    rng = np.random.default_rng(seed)
    two = [GateType.RZZ, GateType.CZ, GateType.CP, GateType.CRz, GateType.CS, GateType.CSdg, GateType.MCZ]
    one = [GateType.Rz, GateType.P, GateType.S, GateType.Sdg, GateType.T, GateType.Tdg, GateType.Z]
    cmds = [Command(GateType.H, q) for q in range(n)]
    for layer in range(layers):
        for _ in range(run_len if run_len is not None else int(rng.integers(4, 3 * n + 4))):
            r = rng.random()
            if r < 0.6 and n >= 2:
                g = two[rng.integers(len(two))]
                a, b = (int(x) for x in rng.choice(n, 2, replace=False))
                cmds.append(Command(g, qubits=[a, b], params=list(rng.uniform(-3, 3, 1 if g in
                                    (GateType.RZZ, GateType.CP, GateType.CRz) else 0))))
            elif r < 0.9:
                g = one[rng.integers(len(one))]
                cmds.append(Command(g, qubits=[int(rng.integers(n))],
                                    params=list(rng.uniform(-3, 3, 1 if g in (GateType.Rz, GateType.P) else 0))))
            elif r < 0.95:
                cmds.append(Command(GateType.GPhase, params=[float(rng.uniform(-3, 3))]))
            else:
                k = 1 if n < 2 else int(rng.integers(1, 3))
                qs = [int(x) for x in rng.choice(n, k, replace=False)]
                cmds.append(Command(GateType.Custom, qubits=qs,
                                    unitary=np.diag(np.exp(1j * rng.uniform(-3, 3, 1 << k)))))
        mixer = [GateType.Rx, GateType.Ry][layer % 2]
        for q in range(n):
            cmds.append(Command(mixer, q, float(rng.uniform(-3, 3))))
    return cmds


@pytest.mark.parametrize("n", [3, 6, 10, 13, 16])
@pytest.mark.parametrize("tile,low", [(0, 0), (6, 2), (9, 3), (12, 4)])
@pytest.mark.parametrize("precision", ["complex128", "complex64"])
def test_phase_runs_merge(n, tile, low, precision):
    """Runs of four or more diagonal phase ops become one merged op per pass
    (reported as a GPhase plus one diagonal per term): the replayed schedule
    equals the circuit, with terms on tile positions below 8, on row
    positions 8.. and outside the tile."""
    cmds = phase_run_circuit(n, 4, 7 * n + tile)
    ref = O.statevector(pack_stream(cmds), n, 0, 0)
    got, sched = scheduled_state(cmds, n, fusion_max_qubits=1, fusion_min_qubits=0, tile_qubits=tile,
                                 low_qubits=low, precision=precision)
    assert np.abs(got - ref).max() < 1e-12
    from openqarp_amd.ir import GateType
    plain = [c for c in cmds if c.gate != GateType.GPhase]
    sched = inspect_schedule(plain, n, fusion_max_qubits=1, fusion_min_qubits=0, tile_qubits=tile, low_qubits=low,
                             precision=precision)
    assert any(c.gate == GateType.GPhase for _, c in sched)   # the merged ops' constants


def test_short_phase_runs_stay_as_they_are():
    """Runs of three diagonal ops are not merged (no GPhase appears)."""
    from openqarp_amd.ir import GateType
    n = 8
    cmds = [c for c in phase_run_circuit(n, 3, 11, run_len=3) if c.gate != GateType.GPhase]
    got, sched = scheduled_state(cmds, n, fusion_max_qubits=1, fusion_min_qubits=0)
    assert np.abs(got - O.statevector(pack_stream(cmds), n, 0, 0)).max() < 1e-12
    assert not any(c.gate == GateType.GPhase for _, c in sched)


def test_qaoa_cost_layers_merge():
    """The QAOA MaxCut circuit at fusion 1: each cost layer's 51 RZZ gates
    (one run) merge, so the schedule holds far fewer items than gates."""
    from openqarp_amd.ir import GateType
    w = W.qaoa_maxcut(16, 4, 5)
    ref = O.statevector(pack_stream(w.commands), 16, 0, 0)
    got, sched = scheduled_state(w.commands, 16, fusion_max_qubits=1, fusion_min_qubits=0, tile_qubits=12,
                                 low_qubits=4)
    assert np.abs(got - ref).max() < 1e-12
    assert sum(c.gate == GateType.GPhase for _, c in sched) >= 4
