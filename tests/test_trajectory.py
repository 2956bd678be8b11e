"""Trajectory path of QarpSimulator::run: mid-circuit Measure, Reset,
classical conditions and branches (qarp_simulator.cpp:1038-1211).

CPU: the oracle's per-shot restatement equals the reference build's counts
and classical registers shot for shot, and the committed fixture; branch
structure errors.  GPU (marked): the shot-batched device path equals the
oracle shot for shot."""
from pathlib import Path

import numpy as np
import pytest

import _oracle as O
from openqarp_amd import workloads as W
from openqarp_amd.ir import Command, GateType
from test_gpu_parity import random_circuit

GOLDEN = Path(__file__).resolve().parent / "golden"


def cond(cmd, bits, values):
    return Command(cmd.gate, qubits=cmd.qubits, params=cmd.params, condition_bits=list(bits),
                   condition_values=[bool(v) for v in values])


def branch(bits, values):
    return cond(Command(GateType.BranchBegin), bits, values)


def circuits(n, seed):
    """Named trajectory circuits on n qubits (n >= 4)."""
    base = W.brickwork_haar(n, 2, seed)
    teleport = base + [
        Command(GateType.Measure, 0, cbits=[0]),
        cond(Command(GateType.X, 1), [0], [1]),
        Command(GateType.H, 1),
        Command(GateType.Measure, 1, cbits=[1]),
        cond(Command(GateType.Rz, 2, params=[0.7]), [0, 1], [1, 0]),
    ] + W.brickwork_haar(n, 1, seed + 1)
    resets = base + [Command(GateType.Reset, 2), Command(GateType.H, 2), Command(GateType.CX, 2, 3),
                     Command(GateType.Measure, 3, cbits=[0]), Command(GateType.Reset, 3),
                     Command(GateType.Measure, 1)] + random_circuit(n, 12, seed)
    branches = base + [
        Command(GateType.Measure, 0, cbits=[0]),
        Command(GateType.Measure, 1, cbits=[1]),
        branch([0], [1]),
        Command(GateType.X, 2), Command(GateType.Measure, 2, cbits=[2]),
        branch([1, 2], [1, 1]), Command(GateType.Ry, 3, params=[0.4]), Command(GateType.BranchEnd),
        Command(GateType.BranchElse),
        Command(GateType.H, 2), Command(GateType.Reset, 0),
        Command(GateType.BranchEnd),
    ] + W.brickwork_haar(n, 1, seed + 2) + [Command(GateType.Measure, q, cbits=[3 + q]) for q in range(n)]
    return {"teleport": teleport, "resets": resets, "branches": branches}


@pytest.mark.skipif(O.reference() is None, reason="reference build absent")
@pytest.mark.parametrize("name", ["teleport", "resets", "branches"])
@pytest.mark.parametrize("n,seed", [(4, 1), (6, 2)])
def test_oracle_trajectory_matches_reference_build(name, n, seed):
    cmds = circuits(n, seed)[name]
    c_ref, h_ref = O.run_trajectory(cmds, n, 300, 1000 + seed, lib="ref")
    c_orc, h_orc = O.run_trajectory(cmds, n, 300, 1000 + seed)
    assert c_orc == c_ref
    assert h_orc == h_ref


def test_oracle_trajectory_matches_fixture():
    d = np.load(GOLDEN / "trajectory6.npz")
    cmds = circuits(int(d["n"]), int(d["seed"]))["branches"]
    counts, hist = O.run_trajectory(cmds, int(d["n"]), int(d["shots"]), int(d["shot_seed"]))
    assert counts == dict(zip(d["keys"].tolist(), d["counts"].tolist()))
    assert np.array_equal(np.array(hist, dtype=bool), d["history"])


def test_branch_structure_errors():
    from openqarp_amd.trajectory import _parse
    with pytest.raises(RuntimeError, match="BranchElse without matching BranchBegin"):
        _parse([Command(GateType.BranchElse)])
    with pytest.raises(RuntimeError, match="BranchBegin without matching BranchEnd"):
        _parse([branch([0], [1]), Command(GateType.X, 0)])


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["teleport", "resets", "branches"])
@pytest.mark.parametrize("n,seed,shots", [(4, 1, 300), (7, 3, 200)])
def test_device_trajectories_match_oracle(name, n, seed, shots):
    """Counts and every shot's classical register equal the oracle's; shots
    run in batches inside one (n + b)-qubit state."""
    from openqarp_amd.simulator import QarpSimulator
    cmds = circuits(n, seed)[name]
    res = QarpSimulator().run(cmds, n, shots, seed=77 + seed)
    counts, hist = O.run_trajectory(cmds, n, shots, 77 + seed)
    assert res.counts == counts
    assert res.cbit_history == hist


@pytest.mark.gpu
def test_device_trajectories_match_fixture_and_batch_split(monkeypatch):
    """The reference's own counts and registers (fixture), with the shots
    split over several batches."""
    from openqarp_amd import trajectory
    from openqarp_amd.simulator import QarpSimulator
    d = np.load(GOLDEN / "trajectory6.npz")
    n = int(d["n"])
    cmds = circuits(n, int(d["seed"]))["branches"]
    monkeypatch.setattr(trajectory, "MAX_BATCH_AMPS", 1 << (n + 4))   # 16 shots per batch
    res = QarpSimulator().run(cmds, n, int(d["shots"]), seed=int(d["shot_seed"]))
    assert res.counts == dict(zip(d["keys"].tolist(), d["counts"].tolist()))
    assert np.array_equal(np.array(res.cbit_history, dtype=bool), d["history"])


@pytest.mark.gpu
def test_wide_shots_take_the_exact_chunk_walk():
    """21-qubit shots (2^21 amplitudes each, 8 per batch): the final draw of
    every shot runs through the terminal sampler's exact chunk scan and
    equals the oracle's sequential walk."""
    from openqarp_amd.simulator import QarpSimulator
    n = 21
    cmds = circuits(n, 4)["teleport"]
    res = QarpSimulator().run(cmds, n, 12, seed=31)
    counts, hist = O.run_trajectory(cmds, n, 12, 31)
    assert res.counts == counts
    assert res.cbit_history == hist
