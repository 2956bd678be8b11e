"""Noise channels on the trajectory path (apply_post_gate_noise /
apply_channel, qarp_simulator.cpp:697-722, device/noise_model.cpp:86-299).

CPU: the per-shot streams (std::mt19937 words, uniforms, libstdc++'s
uniform_int_distribution) against the reference build and the committed
fixtures; the oracle's noisy per-shot restatement against
QarpSimulator(NoiseModel)::run of the reference build, counts and registers
shot for shot, for depolarizing (1q, 2q, many-qubit MCZ), Pauli, bit-flip
and amplitude-damping (Kraus) models composed with +, with conditions and
branches; the builders' checks.  GPU (marked): the shot-batched device path
equals the oracle shot for shot."""
from pathlib import Path

import numpy as np
import pytest

import _oracle as O
from openqarp_amd import workloads as W
from openqarp_amd.ir import Command, GateType
from openqarp_amd.noise import (KrausChannel, NoiseModel, PauliChannel, PauliChannel2q, PauliChannelKq,
                                ShotStreams, make_amplitude_damping, make_bit_flip, make_depolarizing,
                                make_pauli)
from test_trajectory import branch, circuits, cond

GOLDEN = Path(__file__).resolve().parent / "golden"
REF = O.reference() is not None
needs_ref = pytest.mark.skipif(not REF, reason="reference build (oracle/_ref) not present")

G = GateType
ONE_Q = [int(G.H), int(G.X), int(G.Ry), int(G.Rz), int(G.Rx), int(G.S)]
TWO_Q = [int(G.CX), int(G.CZ), int(G.RZZ)]


def models():
    """Named noise models as builder calls (kind, (p...), gates)."""
    return {
        "depolarizing": [(0, (0.08,), ONE_Q + TWO_Q)],
        "pauli_bitflip": [(1, (0.05, 0.02, 0.07), [int(G.H), int(G.Ry)]), (2, (0.1,), [int(G.X), int(G.Rz)])],
        "amplitude_damping": [(3, (0.15,), [int(G.H), int(G.Ry), int(G.Rz), int(G.X)]), (0, (0.05,), TWO_Q)],
        "mcz": [(0, (0.6,), [int(G.MCZ), int(G.CX)])],
    }


def noisy_circuit(n, seed):
    rng = np.random.default_rng(seed)
    cmds = []
    for layer in range(3):
        for q in range(n):
            cmds.append(Command(G.Ry, q, float(rng.uniform(0, 6.3))))
            cmds.append(Command(G.Rz, q, float(rng.uniform(0, 6.3))))
        for q in range(layer % 2, n - 1, 2):
            cmds.append(Command(G.CX, q, q + 1))
    cmds += [Command(G.H, 0), Command(G.MCZ, qubits=list(range(min(n, 4)))), Command(G.RZZ, 1, 2, 0.4),
             Command(G.Measure, 0, cbits=[0]), cond(Command(G.X, 1), [0], [1]), Command(G.CZ, 0, 2),
             branch([0], [1]), Command(G.H, 3), Command(G.BranchElse), Command(G.Ry, 3, 0.3),
             Command(G.BranchEnd), Command(G.Reset, 2), Command(G.H, 2)]
    return cmds


def build_model(entries):
    builders = {0: lambda p, g: make_depolarizing(p[0], g), 1: lambda p, g: make_pauli(p[0], p[1], p[2], g),
                2: lambda p, g: make_bit_flip(p[0], g), 3: lambda p, g: make_amplitude_damping(p[0], g)}
    m = NoiseModel()
    for kind, p, gates in entries:
        m = m + builders[kind](p, [G(x) for x in gates])
    return m


# ── streams ──────────────────────────────────────────────────────────────────

def test_shot_streams_are_std_mt19937():
    f = np.load(GOLDEN / "mt19937.npz")
    s = ShotStreams(5489, 0, 1, 16)
    assert [s.word(0) for _ in range(2000)] == f["words_5489"].tolist()   # across read-ahead extensions
    s = ShotStreams(1000, 234, 3)
    u = np.array([s.uniform(np.array([0]))[0] for _ in range(2000)])   # shot 0: seed 1000 + 234
    assert np.array_equal(u, f["uniforms_1234"])
    assert np.array_equal(O.mt19937(1236, 50), ShotStreams(1236, 0, 1, 60).words[0, :50].astype(np.uint32))


def test_uniform_int_matches_fixture_and_reference():
    f = np.load(GOLDEN / "uniform_int.npz")
    for row, (seed, lo, hi) in enumerate(f["cases"].tolist()):
        want = f["draws"][row]
        s = ShotStreams(int(seed), 0, 1)
        assert [s.uniform_int(0, int(lo), int(hi)) for _ in range(want.size)] == want.tolist()
        assert np.array_equal(O.uniform_ints(int(seed), int(lo), int(hi), want.size), want)
        if REF:
            assert np.array_equal(O.uniform_ints(int(seed), int(lo), int(hi), want.size, lib="ref"), want)


# ── the oracle against the reference build ──────────────────────────────────

@needs_ref
@pytest.mark.parametrize("model", list(models()))
@pytest.mark.parametrize("n,seed", [(4, 1), (5, 2)])
def test_oracle_noisy_trajectories_match_reference_build(model, n, seed):
    cmds = noisy_circuit(n, seed)
    entries = models()[model]
    c_ref, h_ref = O.run_trajectory(cmds, n, 200, 77 + seed, lib="ref", noise=entries)
    c_orc, h_orc = O.run_trajectory(cmds, n, 200, 77 + seed, noise=entries)
    assert c_orc == c_ref
    assert h_orc == h_ref


@needs_ref
def test_disabled_model_is_the_noiseless_run():
    n = 4
    cmds = circuits(n, 3)["teleport"]
    c0, h0 = O.run_trajectory(cmds, n, 150, 5, lib="ref")
    c1, h1 = O.run_trajectory(cmds, n, 150, 5, lib="ref", noise=models()["depolarizing"], enabled=False)
    assert (c0, h0) == (c1, h1)


def test_oracle_matches_noisy_fixture():
    d = np.load(GOLDEN / "noisy_trajectory5.npz")
    cmds = noisy_circuit(int(d["n"]), int(d["seed"]))
    for name in models():
        counts, hist = O.run_trajectory(cmds, int(d["n"]), int(d["shots"]), int(d["shot_seed"]),
                                        noise=models()[name])
        assert counts == dict(zip(d[f"{name}_keys"].tolist(), d[f"{name}_counts"].tolist()))
        assert np.array_equal(np.array(hist, dtype=bool), d[f"{name}_history"])


# ── the model and its builders ───────────────────────────────────────────────

def test_builders_and_composition():
    m = make_depolarizing(0.3, [G.H, G.CX, G.MCZ])
    assert m.channel_for(Command(G.H, 0)) == PauliChannel(0.3 / 3.0, 0.3 / 3.0, 0.3 / 3.0)
    assert m.channel_for(Command(G.CX, 0, 1)) == PauliChannel2q((0.3 / 15,) * 15)
    assert m.channel_for(Command(G.MCZ, qubits=[0, 1, 2])) == PauliChannelKq(0.3)
    assert m.channel_for(Command(G.Measure, 0)) is None and m.channel_for(Command(G.Rz, 0, 0.1)) is None
    both = m + make_bit_flip(0.2, [G.H])
    assert both.channel_for(Command(G.H, 0)) == PauliChannel(0.2, 0.0, 0.0)   # right operand wins
    off = make_bit_flip(0.2, [G.X])
    off.enabled = False
    assert not (both + off).enabled and not (both + off).active() and both.active()
    k = make_amplitude_damping(0.36, [G.X]).channel_for(Command(G.X, 0))
    assert isinstance(k, KrausChannel) and k.K[0][1, 1] == np.sqrt(1.0 - 0.36) and k.K[1][0, 1] == np.sqrt(0.36)
    assert not NoiseModel().active() and not NoiseModel().has_any_channel()
    with pytest.raises(ValueError, match=r"make_depolarizing: p must lie in \[0, 1\]"):
        make_depolarizing(1.5, [G.H])
    with pytest.raises(ValueError, match="unsupported qubit count"):
        make_depolarizing(0.1, [G.CCX])
    with pytest.raises(ValueError, match="non-negative and sum"):
        make_pauli(0.5, 0.4, 0.3, [G.H])
    with pytest.raises(ValueError, match="is not 1-qubit"):
        make_amplitude_damping(0.1, [G.CX])


def test_noisy_simulator_refuses_statevectors():
    from openqarp_amd.simulator import QarpSimulator
    sim = QarpSimulator(make_depolarizing(0.1, [G.H]))
    assert sim.noise_active()
    with pytest.raises(RuntimeError, match="noise model is active"):
        sim.statevector([Command(G.H, 0)], 1)
    sim.noise_model.enabled = False
    assert not sim.noise_active()


# ── the device ───────────────────────────────────────────────────────────────

@pytest.mark.gpu
@pytest.mark.parametrize("model", list(models()))
@pytest.mark.parametrize("n,seed", [(4, 1), (7, 3)])
def test_device_noisy_trajectories_match_oracle(model, n, seed):
    from openqarp_amd.simulator import QarpSimulator
    cmds = noisy_circuit(n, seed)
    entries = models()[model]
    sim = QarpSimulator(build_model(entries))
    shots = 300
    res = sim.run(cmds, n, shots, seed=900 + seed)
    counts, hist = O.run_trajectory(cmds, n, shots, 900 + seed, noise=entries)
    assert res.counts == counts
    assert res.cbit_history == hist


@pytest.mark.gpu
def test_device_noisy_fixture_and_many_batches():
    """The committed reference fixture, and a run spanning several shot
    batches (shots beyond one batch draw from streams base + shot)."""
    from openqarp_amd import trajectory
    from openqarp_amd.simulator import QarpSimulator
    d = np.load(GOLDEN / "noisy_trajectory5.npz")
    n = int(d["n"])
    cmds = noisy_circuit(n, int(d["seed"]))
    for name in models():
        res = QarpSimulator(build_model(models()[name])).run(cmds, n, int(d["shots"]), seed=int(d["shot_seed"]))
        assert res.counts == dict(zip(d[f"{name}_keys"].tolist(), d[f"{name}_counts"].tolist()))
        assert np.array_equal(np.array(res.cbit_history, dtype=bool), d[f"{name}_history"])
    old = trajectory.MAX_BATCH_AMPS
    try:
        trajectory.MAX_BATCH_AMPS = 1 << (n + 5)      # batches of 32 shots
        res = QarpSimulator(build_model(models()["amplitude_damping"])).run(cmds, n, 150, seed=3)
    finally:
        trajectory.MAX_BATCH_AMPS = old
    counts, hist = O.run_trajectory(cmds, n, 150, 3, noise=models()["amplitude_damping"])
    assert res.counts == counts and res.cbit_history == hist
