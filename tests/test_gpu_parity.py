"""Parity of the HIP backend with the oracle and the reference's golden
vectors.  Tolerances: 1e-12 absolute for complex128 amplitudes and <H>,
bit-exact sampled outcomes, 1e-10 for adjoint gradients against the oracle
(1e-8 against finite differences is the bar the backend targets)."""
import math

import numpy as np
import pytest

import _fixtures as F
import _oracle as O
from openqarp_amd import workloads as W
from openqarp_amd._abi import pack_observable, pack_stream
from openqarp_amd.errors import CapabilityError
from openqarp_amd.ir import Command, GateType, Param
from openqarp_amd.simulator import DeviceState, QarpSimulator

pytestmark = pytest.mark.gpu

ATOL = 1e-12


def gpu_sim(**kw):
    s = QarpSimulator(**kw)
    return s


# ── statevector ────────────────────────────────────────────────────────────

@pytest.mark.parametrize("name", ["hea12.npz", "hea16.npz", "hea8_unfused.npz"])
def test_statevector_matches_reference_fixture(name):
    f = F.load(name)
    n = int(f["n"])
    w = W.c1_hea(n, int(f["layers"]), int(f["seed"]))
    sim = gpu_sim()
    sim.fusion_max_qubits, sim.fusion_min_qubits = (int(x) for x in f["fusion"])
    sv = sim.statevector(w.commands, n)
    assert np.abs(sv - f["state"]).max() < ATOL


@pytest.mark.parametrize("k", [2, 3, 4, 5])
def test_brickwork_every_fusion_width(k):
    f = F.load("brickwork10.npz")
    n = int(f["n"])
    cmds = W.brickwork_haar(n, int(f["layers"]), int(f["seed"]))
    sim = gpu_sim()
    sim.fusion_max_qubits, sim.fusion_min_qubits = k, 0
    assert np.abs(sim.statevector(cmds, n) - f[f"state_k{k}"]).max() < ATOL


def test_fused_stream_matches_reference_fixture():
    f = F.load("brickwork10.npz")
    n = int(f["n"])
    cmds = W.brickwork_haar(n, int(f["layers"]), int(f["seed"]))
    sim = gpu_sim()
    sim.fusion_max_qubits, sim.fusion_min_qubits = 5, 0
    ours = sim.fused_stream(cmds, n)
    assert len(ours) == len(f["fused_gates"])
    for i, (g, qs, _, m) in enumerate(ours):
        assert g == int(f["fused_gates"][i])
        assert qs == tuple(int(q) for q in f["fused_qubits"][i] if q >= 0)
        if m is not None:
            d = m.shape[0]
            assert np.abs(m - f["fused_mats"][i][:d, :d]).max() < ATOL


ALL_GATES = [GateType.X, GateType.Y, GateType.Z, GateType.H, GateType.S, GateType.Sdg, GateType.T,
             GateType.Tdg, GateType.SX, GateType.SXdg, GateType.Id, GateType.Rx, GateType.Ry, GateType.Rz,
             GateType.P, GateType.U, GateType.CX, GateType.CY, GateType.CZ, GateType.SWAP, GateType.ECR,
             GateType.iSWAP, GateType.iSWAPdg, GateType.CH, GateType.CS, GateType.CSdg, GateType.CSX,
             GateType.CSXdg, GateType.CRx, GateType.CRy, GateType.CRz, GateType.CP, GateType.RZZ,
             GateType.RXX, GateType.RYY, GateType.CU, GateType.CCX, GateType.CSWAP, GateType.MCZ,
             GateType.GPhase, GateType.Barrier]
N_PARAMS = {GateType.Rx: 1, GateType.Ry: 1, GateType.Rz: 1, GateType.P: 1, GateType.U: 3, GateType.CRx: 1,
            GateType.CRy: 1, GateType.CRz: 1, GateType.CP: 1, GateType.RZZ: 1, GateType.RXX: 1,
            GateType.RYY: 1, GateType.CU: 4, GateType.GPhase: 1}


def random_circuit(n, n_gates, seed, with_custom=True):
    rng = np.random.default_rng(seed)
    cmds = []
    for _ in range(n_gates):
        g = ALL_GATES[rng.integers(len(ALL_GATES))]
        if g in (GateType.CCX, GateType.CSWAP):
            nq = 3
        elif g == GateType.MCZ:
            if n < 2:
                continue
            nq = int(rng.integers(2, min(n, 6) + 1))
        elif g == GateType.GPhase:
            nq = 0
        elif g == GateType.Barrier:
            nq = int(rng.integers(0, 3))
        elif g.value >= 16:
            nq = 2
        else:
            nq = 1
        if nq > n:
            continue
        qs = [int(q) for q in rng.choice(n, nq, replace=False)]
        cmds.append(Command(g, qubits=qs, params=list(rng.uniform(-3, 3, N_PARAMS.get(g, 0)))))
        if with_custom and rng.random() < 0.1 and n >= 2:
            k = int(rng.integers(1, min(n, 3) + 1))
            cmds.append(Command(GateType.Custom, qubits=[int(q) for q in rng.choice(n, k, replace=False)],
                                unitary=W.haar_unitary(1 << k, rng)))
    return cmds


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 11, 13])
@pytest.mark.parametrize("fusion", [(0, 0), (1, 0), (3, 0), (5, 0), (3, 12)])
def test_random_circuits_every_gate(n, fusion):
    cmds = random_circuit(n, 80, 1000 * n + fusion[0])
    sim = gpu_sim()
    sim.fusion_max_qubits, sim.fusion_min_qubits = fusion
    ref = O.statevector(pack_stream(cmds), n, *fusion)
    assert np.abs(sim.statevector(cmds, n) - ref).max() < ATOL


@pytest.mark.parametrize("tile,low", [(5, 2), (6, 3), (8, 4), (10, 5), (12, 4)])
def test_multi_pass_schedules(tile, low):
    """Small tiles force many passes with scattered tile qubits."""
    n = 14
    cmds = W.brickwork_haar(n, 8, 3) + random_circuit(n, 60, 7)
    sim = gpu_sim(tile_qubits=tile, low_qubits=low)
    sim.fusion_max_qubits, sim.fusion_min_qubits = 4, 0
    ref = O.statevector(pack_stream(cmds), n, 4, 0)
    out = sim.statevector(cmds, n)
    assert sim.last_stats["passes"] >= 1
    assert np.abs(out - ref).max() < ATOL


@pytest.mark.parametrize("k", [3, 4, 5])
def test_matrix_core_path_matches_oracle(k):
    """Dense k = 3..5 ops run on the matrix cores in complex128; the VALU
    path (use_mfma = 0) and the oracle must agree with them."""
    n = 15
    cmds = W.brickwork_haar(n, 10, 40 + k)
    ref = O.statevector(pack_stream(cmds), n, k, 0)
    on = gpu_sim()
    on.fusion_max_qubits, on.fusion_min_qubits = k, 0
    got = on.statevector(cmds, n)
    assert on.last_stats["mfma_ops"] > 0
    assert np.abs(got - ref).max() < ATOL
    off = gpu_sim(use_mfma=0)
    off.fusion_max_qubits, off.fusion_min_qubits = k, 0
    got_off = off.statevector(cmds, n)
    assert off.last_stats["mfma_ops"] == 0
    assert np.abs(got_off - ref).max() < ATOL


@pytest.mark.parametrize("k", [3, 4, 5])
def test_complex64_matrix_core_path(k):
    """complex64 dense k = 3..5 on the FP32 matrix cores (the result-row
    layout differs from FP64's) against the oracle and the VALU path."""
    n = 15
    cmds = W.brickwork_haar(n, 10, 60 + k)
    ref = O.statevector(pack_stream(cmds), n, k, 0)
    on = gpu_sim(precision="complex64")
    on.fusion_max_qubits, on.fusion_min_qubits = k, 0
    got = on.statevector(cmds, n)
    assert on.last_stats["mfma_ops"] > 0
    assert np.abs(got - ref).max() < 2e-5
    off = gpu_sim(precision="complex64", use_mfma=0)
    off.fusion_max_qubits, off.fusion_min_qubits = k, 0
    got_off = off.statevector(cmds, n)
    assert off.last_stats["mfma_ops"] == 0
    assert np.abs(got_off - got).max() < 2e-5


@pytest.mark.parametrize("fusion,tile", [(2, 11), (2, 12), (1, 11)])
def test_register_blocks_at_22_qubits(fusion, tile):
    """2^22 amplitudes in 2^11 / 2^12 tiles, passes of register blocks and
    dense k <= 2 ops."""
    n = 22
    cmds = W.brickwork_haar(n, 12, 9) + random_circuit(n, 40, 5)
    ref = O.statevector(pack_stream(cmds), n, fusion, 0)
    sim = gpu_sim(tile_qubits=tile)
    sim.fusion_max_qubits, sim.fusion_min_qubits = fusion, 0
    assert np.abs(sim.statevector(cmds, n) - ref).max() < ATOL


@pytest.mark.parametrize("precision,tile", [("complex128", 12), ("complex64", 13)])
def test_largest_tiles(precision, tile):
    """64-KiB tiles (2^12 complex128, 2^13 complex64 amplitudes): dense
    k = 1..5 (matrix cores on and off in complex128), register blocks, the
    adjoint's two-state passes, and a batched step whose small instances
    share the largest tile array."""
    n = 16
    tol = ATOL if precision == "complex128" else 2e-5
    cmds = W.brickwork_haar(n, 10, 17) + random_circuit(n, 50, 3)
    for k in (1, 2, 3, 4, 5):
        for mfma in ((-1, 0) if precision == "complex128" else (-1,)):
            sim = gpu_sim(tile_qubits=tile, precision=precision, use_mfma=mfma)
            sim.fusion_max_qubits, sim.fusion_min_qubits = k, 0
            ref = O.statevector(pack_stream(cmds), n, k, 0)
            assert np.abs(sim.statevector(cmds, n) - ref).max() < tol, (k, mfma)
    w = W.c3_gradient(14, 3, 6, 40)
    sim = gpu_sim(tile_qubits=min(tile, 14), precision=precision)
    g = np.array(sim.run_gradient(w.commands, 14, w.observable, w.values, w.symbols))
    idx = {s: i for i, s in enumerate(w.symbols)}
    ref = O.run_gradient(pack_stream(w.commands, w.values, idx), 14, pack_observable(w.observable), len(idx))
    assert np.abs(g - ref).max() < (1e-10 if precision == "complex128" else 1e-4)
    inst = [(b.commands, b.n_qubits, b.observable, b.values, b.symbols) for b in W.batched_vqe(10, 2, 5, 14, 1, 4)]
    sim = gpu_sim(tile_qubits=tile, precision=precision)
    e, gr = sim.batch_energy_grad(inst)
    for i, (c, m, o, v, s) in enumerate(inst):
        ix = {q: j for j, q in enumerate(s)}
        ps = pack_stream(c, values=v, param_index=ix)
        sv = O.statevector(ps, m, 3, 12)
        po = pack_observable(o)
        scale = sum(abs(complex(x)) for _, x in o)
        te, tg = (1e-12, 1e-10) if precision == "complex128" else (2e-5 * scale, 2e-5 * scale)
        assert abs(e[i] - O.transition(sv, sv, m, po).real) < te
        assert np.abs(gr[i] - O.run_gradient(ps, m, po, len(s))).max() < tg


@pytest.mark.parametrize("precision", ["complex128", "complex64"])
def test_relabeling_store_permutations(precision):
    """Passes whose stores permute tile positions (moving the next pass's
    qubits into the contiguous low positions) against the same circuits
    without relabeling and the oracle: one-state passes at fusion widths
    2..5, the adjoint's two-state passes and a batched step."""
    from openqarp_amd.simulator import inspect_schedule
    n, tile, low = 14, 9, 3
    tol = ATOL if precision == "complex128" else 2e-5
    cmds = W.brickwork_haar(n, 8, 3) + random_circuit(n, 40, 9)
    for k in (2, 3, 4, 5):
        sched = inspect_schedule(cmds, n, fusion_max_qubits=k, fusion_min_qubits=0, tile_qubits=tile,
                                 low_qubits=low, precision=precision)
        swap = np.eye(4)[[0, 2, 1, 3]]
        assert any(c.unitary is not None and c.unitary.shape == (4, 4) and np.array_equal(c.unitary, swap)
                   for _, c in sched), k
        ref = O.statevector(pack_stream(cmds), n, k, 0)
        for relabel in (-1, 0):
            sim = gpu_sim(tile_qubits=tile, low_qubits=low, precision=precision, relabel_qubits=relabel)
            sim.fusion_max_qubits, sim.fusion_min_qubits = k, 0
            assert np.abs(sim.statevector(cmds, n) - ref).max() < tol, (k, relabel)
    w = W.c3_gradient(14, 3, 6, 40)
    idx = {s: i for i, s in enumerate(w.symbols)}
    ref = O.run_gradient(pack_stream(w.commands, w.values, idx), 14, pack_observable(w.observable), len(idx))
    sim = gpu_sim(tile_qubits=8, low_qubits=low, precision=precision)
    g = np.array(sim.run_gradient(w.commands, 14, w.observable, w.values, w.symbols))
    assert np.abs(g - ref).max() < (1e-10 if precision == "complex128" else 1e-4)
    inst = [(b.commands, b.n_qubits, b.observable, b.values, b.symbols) for b in W.batched_vqe(12, 2, 5, 14, 1, 3)]
    sim = gpu_sim(tile_qubits=8, low_qubits=low, precision=precision)
    e, gr = sim.batch_energy_grad(inst)
    for i, (c, m, o, v, s) in enumerate(inst):
        ix = {q: j for j, q in enumerate(s)}
        ps = pack_stream(c, values=v, param_index=ix)
        sv = O.statevector(ps, m, 3, 12)
        po = pack_observable(o)
        scale = sum(abs(complex(x)) for _, x in o)
        te, tg = (1e-12, 1e-10) if precision == "complex128" else (2e-5 * scale, 2e-5 * scale)
        assert abs(e[i] - O.transition(sv, sv, m, po).real) < te
        assert np.abs(gr[i] - O.run_gradient(ps, m, po, len(s))).max() < tg


def test_wide_fusion_blocks_6_to_8():
    n = 11
    cmds = W.brickwork_haar(n, 8, 4)
    for k in (6, 7, 8):
        sim = gpu_sim()
        sim.fusion_max_qubits, sim.fusion_min_qubits = k, 0
        ref = O.statevector(pack_stream(cmds), n, k, 0)
        assert np.abs(sim.statevector(cmds, n) - ref).max() < ATOL


def test_initial_state_and_validation():
    n = 9
    rng = np.random.default_rng(2)
    psi = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    psi /= np.linalg.norm(psi)
    cmds = random_circuit(n, 40, 3)
    sim = gpu_sim()
    ref = O.statevector(pack_stream(cmds), n, 3, 12, init=psi)
    assert np.abs(sim.statevector(cmds, n, initial_state=psi) - ref).max() < ATOL
    with pytest.raises(ValueError, match="not normalised"):
        sim.statevector(cmds, n, initial_state=psi * 1.001)
    with pytest.raises(ValueError, match="initial_state length"):
        sim.statevector(cmds, n, initial_state=psi[:-1])


def test_statevector_rejects_trajectory_and_bad_qubits():
    sim = gpu_sim()
    with pytest.raises(RuntimeError, match="not deterministic"):
        sim.statevector([Command(GateType.H, 0), Command(GateType.Measure, 0, cbits=[0]),
                         Command(GateType.X, 0)], 1)
    with pytest.raises(RuntimeError, match="not deterministic"):
        sim.statevector([Command(GateType.Reset, 0)], 1)
    with pytest.raises(ValueError, match="out of range"):
        sim.statevector([Command(GateType.H, 3)], 2)
    with pytest.raises(ValueError, match="exceeds the maximum"):
        sim.fusion_max_qubits = 9


def test_complex64_state_error_against_complex128():
    n = 14
    w = W.c1_hea(n, 6, 3)
    ref = O.statevector(pack_stream(w.commands), n, 3, 12)
    s64 = QarpSimulator(precision="complex64")
    out = s64.statevector(w.commands, n)
    err = np.abs(out - ref).max()
    assert err < 1e-5
    e64 = s64.expectation(w.commands, n, w.observable)
    e128 = O.transition(ref, ref, n, pack_observable(w.observable)).real
    assert abs(e64 - e128) < 1e-4


def test_norm_is_preserved_at_24_qubits():
    n = 24
    sim = gpu_sim()
    st = sim.new_state(n)
    sim.fusion_max_qubits = 4
    sim.apply(W.brickwork_haar(n, 12, 1), st)
    assert abs(st.norm2() - 1.0) < 1e-12


def test_circuit_then_inverse_returns_to_zero_state_at_22_qubits():
    n = 22
    cmds = W.brickwork_haar(n, 10, 8) + random_circuit(n, 120, 9)
    inv = [c.dagger() for c in reversed(cmds)]
    sim = gpu_sim()
    st = sim.new_state(n)
    sim.apply(cmds, st)
    sim.apply(inv, st)
    out = st.download()
    assert abs(out[0] - 1.0) < 1e-11
    assert np.abs(out[1:]).max() < 1e-11


# ── Pauli expectation ──────────────────────────────────────────────────────

def test_transition_matches_reference_fixture():
    f = F.load("transition8.npz")
    n = int(f["n"])
    v = gpu_sim().transition(f["bra"], f["ket"], n, F.observable(f))
    assert abs(v - complex(f["value"])) < ATOL


@pytest.mark.parametrize("n", [0, 1, 2, 5, 6, 7, 12, 13, 16])
def test_transition_random_sizes(n):
    rng = np.random.default_rng(n + 40)
    dim = 1 << n
    bra = rng.standard_normal(dim) + 1j * rng.standard_normal(dim)
    ket = rng.standard_normal(dim) + 1j * rng.standard_normal(dim)
    bra /= np.linalg.norm(bra)
    ket /= np.linalg.norm(ket)
    obs = W.random_pauli_sum(max(n, 1), 30, n, probs=(0.4, 0.2, 0.2, 0.2)) if n else [([], 0.7 - 0.2j)]
    obs = [([(q, l) for q, l in p if q < n], c * (1 + 0.5j)) for p, c in obs]
    got = gpu_sim().transition(bra, ket, n, obs)
    ref = O.transition(bra, ket, n, pack_observable(obs))
    assert abs(got - ref) < ATOL


@pytest.mark.parametrize("precision", ["complex128", "complex64"])
def test_hermitian_half_sum_expectation(precision):
    """<psi|H|psi> with real coefficients sums half the index pairs per flip
    mask (twice the real part): masks whose top bit falls in a lane bit, a
    row bit and a bit above the 2^12 chunk, plus Z-only terms, all against the
    oracle; complex coefficients take the full sum."""
    n = 15
    rng = np.random.default_rng(9)
    sim = gpu_sim(precision=precision)
    st = sim.new_state(n)
    sim.apply(W.brickwork_haar(n, 5, 2), st)
    amps = st.download().astype(np.complex128)
    tol = ATOL if precision == "complex128" else 1e-6
    obs = []
    for top in (1, 2, 5, 7, 9, 11, 12, 14):
        letters = [(top, "XY"[top % 2])] + [(int(q), "XYZ"[int(rng.integers(3))])
                                             for q in rng.choice(top, min(top, 3), replace=False)]
        obs.append((letters, float(rng.standard_normal())))
    obs += [([(3, "Z"), (13, "Z")], 0.3), ([], -0.25)]
    got = sim.expectation_state(st, obs)
    ref = O.transition(amps, amps, n, pack_observable(obs))
    assert abs(got - ref.real) < tol and abs(ref.imag) < 1e-12
    other = sim.new_state(n)
    other.upload(amps)
    assert abs(sim.transition_state(other, st, obs) - ref) < tol      # two states: the full sum
    cobs = [(p, c * (1 + 0.25j)) for p, c in obs]
    want = O.transition(amps, amps, n, pack_observable(cobs))
    assert abs(sim.transition_state(st, st, cobs) - want) < tol       # complex coefficients: the full sum


def test_expectation_matches_reference_fixture_energy():
    f = F.load("hea16.npz")
    n = int(f["n"])
    w = W.c1_hea(n, int(f["layers"]), int(f["seed"]))
    assert abs(gpu_sim().expectation(w.commands, n, F.observable(f)) - float(f["energy"])) < ATOL


def test_h2_fci_known_answer():
    from test_oracle import h2_jordan_wigner
    obs = h2_jordan_wigner()
    w, v = np.linalg.eigh(F.pauli_matrix(obs, 4))
    sim = gpu_sim()
    e = sim.transition(v[:, 0], v[:, 0], 4, obs)
    assert e.real == pytest.approx(-1.1373060357534004, abs=ATOL)


def test_transition_errors():
    sim = gpu_sim()
    sv = np.zeros(4, np.complex128)
    sv[0] = 1
    with pytest.raises(ValueError, match="out of range"):
        sim.transition(sv, sv, 2, [([(2, "X")], 1.0)])
    with pytest.raises(ValueError, match="repeated"):
        sim.transition(sv, sv, 2, [([(0, "X"), (0, "Z")], 1.0)])
    with pytest.raises(ValueError, match="invalid Pauli"):
        sim.transition(sv, sv, 2, [([(0, "Q")], 1.0)])
    with pytest.raises(ValueError, match=r"must be in \[0, 62\]"):
        sim.transition(sv, sv, 63, [])
    assert sim.transition(sv, sv, 2, []) == 0


def test_apply_observable_matches_oracle():
    from openqarp_amd import _lib
    import ctypes as C
    f = F.load("transition8.npz")
    n = int(f["n"])
    obs = F.observable(f)
    a = DeviceState(n).upload(f["ket"])
    b = DeviceState(n)
    po = pack_observable(obs)
    _lib.check(_lib.load().qhip_apply_observable(a.handle, po.ref, b.handle))
    assert np.abs(b.download() - O.apply_observable(f["ket"], n, po)).max() < ATOL


@pytest.mark.parametrize("precision", ["complex128", "complex64"])
@pytest.mark.parametrize("n,n_terms,probs", [(12, 40, (0.4, 0.2, 0.2, 0.2)), (13, 60, (0.5, 0.2, 0.2, 0.1)),
                                             (15, 200, (0.4, 0.2, 0.2, 0.2)), (17, 150, (0.6, 0.15, 0.15, 0.1))])
def test_apply_observable_chunked_matches_oracle(precision, n, n_terms, probs):
    """H|psi> on states of 2^12 amplitudes or more: one workgroup per 2^12
    chunk, one staged partner chunk per run of flip masks with equal bits 12
    and up (the next run's chunk loaded while the current one computes).
    Random sums with X/Y on every qubit give runs of one group and of many,
    a diagonal run and complex coefficients; the oracle works from the
    amplitudes the device holds (rounded once for complex64)."""
    from openqarp_amd import _lib
    rng = np.random.default_rng(n * 7 + n_terms)
    dim = 1 << n
    psi = rng.standard_normal(dim) + 1j * rng.standard_normal(dim)
    psi /= np.linalg.norm(psi)
    obs = W.random_pauli_sum(n, n_terms, n + 3, probs=probs)
    obs = [(p, c * (1 + 0.5j)) if i % 3 == 0 else (p, c) for i, (p, c) in enumerate(obs)]
    obs += [([(n - 1, "X")], 0.75), ([(0, "Z"), (n - 1, "Z")], -0.5)]
    po = pack_observable(obs)
    a = DeviceState(n, precision).upload(psi)
    b = DeviceState(n, precision)
    _lib.check(_lib.load().qhip_apply_observable(a.handle, po.ref, b.handle))
    ref = O.apply_observable(a.download(), n, po)
    got = b.download()
    scale = max(1.0, float(np.abs(ref).max()))
    tol = ATOL if precision == "complex128" else 1e-6   # complex64: the output's one rounding
    assert np.abs(got - ref).max() < tol * scale
    one = pack_observable(obs[:1])                      # a single group: one run
    b2 = DeviceState(n, precision)
    _lib.check(_lib.load().qhip_apply_observable(a.handle, one.ref, b2.handle))
    assert np.abs(b2.download() - O.apply_observable(a.download(), n, one)).max() < tol * scale


# ── sampling ───────────────────────────────────────────────────────────────

@pytest.mark.parametrize("name", ["sample10.npz", "hea12.npz", "hea16.npz"])
def test_samples_bit_exact_against_reference_fixture(name):
    f = F.load(name)
    n = int(f["n"])
    sim = gpu_sim()
    st = DeviceState(n).upload(f["state"])
    out = sim.sample_state(st, int(f["shots"]), int(f["shot_seed"]))
    assert (out == f["samples"]).all()


@pytest.mark.parametrize("n", [0, 1, 3, 11, 12, 20])
def test_samples_bit_exact_against_oracle(n):
    rng = np.random.default_rng(n)
    amp = rng.standard_normal(1 << n) * rng.exponential(1.0, 1 << n) + 1j * rng.standard_normal(1 << n)
    amp[rng.random(1 << n) < 0.1] = 0
    if not amp.any():
        amp[0] = 1
    amp /= np.linalg.norm(amp)
    st = DeviceState(n).upload(amp)
    got = gpu_sim().sample_state(st, 20000, 123 + n)
    assert (got == O.sample(amp, n, 20000, 123 + n)).all()


def test_uniforms_at_every_cdf_boundary():
    """Draws placed exactly on (and one ulp either side of) CDF entries."""
    n = 12
    rng = np.random.default_rng(77)
    amp = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    amp /= np.linalg.norm(amp)
    cdf = O.cdf(amp, n)
    pick = cdf[rng.integers(0, (1 << n) - 1, 3000)]
    u = np.concatenate([pick, np.nextafter(pick, 0), np.nextafter(pick, 1), [0.0]])
    u = u[(u >= 0) & (u < 1)]
    st = DeviceState(n).upload(amp)
    assert (gpu_sim().sample_uniforms(st, u) == O.sample_uniforms(amp, n, u)).all()


def test_batch_run_seeds_each_set_with_seed_plus_index():
    """batch_run substitutes each parameter set and samples it with seed + i
    in uint32 arithmetic (qarp_simulator.cpp:1257-1266): the counts equal the
    oracle's sampler on the oracle's statevector with that seed, across the
    2^32 wrap."""
    n = 9
    a, b = Param.symbol("a"), Param.symbol("b")
    cmds = W.brickwork_haar(n, 3, 5) + [Command(GateType.Rx, 2, params=[a]), Command(GateType.Ry, 6, params=[b]),
                                        Command(GateType.CX, 2, 6)]
    sets = [{"a": 0.3, "b": -1.1}, {"a": 2.0, "b": 0.4}, {"a": -0.7, "b": 3.0}]
    seed, shots = 0xFFFFFFFE, 3000
    res = gpu_sim().batch_run(cmds + [Command(GateType.Measure, q, cbits=[q]) for q in range(n)], n, shots,
                              sets, seed=seed)
    assert len(res) == len(sets)
    for i, (r, ps) in enumerate(zip(res, sets)):
        sv = O.statevector(pack_stream([c.substitute(ps) for c in cmds]), n, 3, 12)
        vals, cnt = np.unique(O.sample(sv, n, shots, (seed + i) & 0xFFFFFFFF), return_counts=True)
        assert r.counts == {int(v): int(c) for v, c in zip(vals, cnt)}, i


def test_run_counts_and_cbit_history():
    f = F.load("hea12.npz")
    n = int(f["n"])
    w = W.c1_hea(n, int(f["layers"]), int(f["seed"]))
    cmds = w.commands + [Command(GateType.Measure, q, cbits=[q]) for q in range(n)]
    res = gpu_sim().run(cmds, n, int(f["shots"]), int(f["shot_seed"]))
    vals, cnt = np.unique(f["samples"], return_counts=True)
    assert res.counts == {int(v): int(c) for v, c in zip(vals, cnt)}
    assert res.n_cbits == n
    first = f["samples"][0]
    assert res.cbit_history[0] == [bool((int(first) >> q) & 1) for q in range(n)]


def test_sampling_at_22_qubits_matches_oracle():
    n = 22
    sim = gpu_sim()
    st = sim.new_state(n)
    sim.apply(W.brickwork_haar(n, 6, 2), st)
    sv = st.download()
    assert (sim.sample_state(st, 4000, 5) == O.sample(sv, n, 4000, 5)).all()


# ── adjoint gradient ───────────────────────────────────────────────────────

def test_gradient_matches_reference_fixture():
    f = F.load("gradient8.npz")
    n = int(f["n"])
    w = W.c3_gradient(n, int(f["layers"]), int(f["seed"]), int(f["n_terms"]))
    g, e = gpu_sim().run_gradient_energy(w.commands, n, w.observable, w.values, w.symbols)
    assert np.abs(g - f["grad"]).max() < 1e-10
    assert abs(e - float(f["energy"])) < ATOL


def test_qaoa_gradient_matches_reference_fixture():
    f = F.load("qaoa10.npz")
    n = int(f["n"])
    w = W.qaoa_maxcut(n, int(f["p"]), int(f["seed"]), symbolic=True)
    g = gpu_sim().run_gradient(w.commands, n, w.observable, w.values, w.symbols)
    assert np.abs(np.array(g) - f["grad"]).max() < 1e-10


def test_gradient_every_generator_rule_against_oracle():
    from test_oracle import _fd_energy
    n = 3
    t = [Param.symbol(f"t{i}") for i in range(12)]
    cmds = [Command(GateType.H, 0), Command(GateType.H, 1), Command(GateType.H, 2),
            Command(GateType.Rx, 0, t[0]), Command(GateType.Ry, 1, t[1] * 2.0 + 0.3),
            Command(GateType.Rz, 2, t[2]), Command(GateType.P, 0, t[3]),
            Command(GateType.CRx, 0, 1, t[4]), Command(GateType.CRy, 1, 2, t[5]),
            Command(GateType.CRz, 2, 0, t[6]), Command(GateType.CP, 0, 2, t[7]),
            Command(GateType.RXX, 0, 1, t[8]), Command(GateType.RYY, 1, 2, t[9]),
            Command(GateType.RZZ, 0, 2, t[10] - t[0]), Command(GateType.GPhase, params=[t[11]]),
            Command(GateType.S, 1), Command(GateType.SX, 2), Command(GateType.iSWAP, 0, 1)]
    names = [f"t{i}" for i in range(12)]
    vals = dict(zip(names, np.random.default_rng(4).uniform(-2, 2, 12)))
    obs = [([(0, "X"), (1, "Y")], 0.7), ([(2, "Z")], -0.4), ([(0, "Y"), (2, "X")], 0.3)]
    g = np.array(gpu_sim().run_gradient(cmds, n, obs, vals, names))
    ref = O.run_gradient(pack_stream(cmds, vals, {s: i for i, s in enumerate(names)}), n,
                         pack_observable(obs), len(names))
    assert np.abs(g - ref).max() < 1e-10
    w = W.Workload(n, cmds, obs, vals, names)
    assert np.abs(g - _fd_energy(w, n)).max() < 1e-7


@pytest.mark.parametrize("tile", [0, 5, 7])
def test_gradient_multi_pass_against_oracle(tile):
    n = 12
    w = W.c3_gradient(n, 3, 6, 60)
    sim = gpu_sim(tile_qubits=tile)
    g = np.array(sim.run_gradient(w.commands, n, w.observable, w.values, w.symbols))
    idx = {s: i for i, s in enumerate(w.symbols)}
    ref = O.run_gradient(pack_stream(w.commands, w.values, idx), n, pack_observable(w.observable), len(idx))
    assert np.abs(g - ref).max() < 1e-10


def test_gradient_phi_against_oracle():
    n = 7
    w = W.c3_gradient(n, 2, 12, 10)
    rng = np.random.default_rng(1)
    phi = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    idx = {s: i for i, s in enumerate(w.symbols)}
    g = np.array(gpu_sim().run_gradient_phi(w.commands, n, phi, w.values, w.symbols))
    ref = O.run_gradient_phi(pack_stream(w.commands, w.values, idx), n, phi, len(idx))
    assert np.abs(g - ref).max() < 1e-10


def test_gradient_capability_and_trajectory_errors():
    sim = gpu_sim()
    cmds = [Command(GateType.U, 0, params=[Param.symbol("a"), 0.1, 0.2])]
    with pytest.raises(CapabilityError):
        sim.run_gradient(cmds, 1, [([(0, "Z")], 1.0)], {"a": 0.3}, ["a"])
    with pytest.raises(RuntimeError, match="not differentiable"):
        sim.run_gradient([Command(GateType.Reset, 0)], 1, [([(0, "Z")], 1.0)], {}, [])


def test_gphase_gradient_known_answer():
    from test_oracle import gphase_circuit
    g = gpu_sim().run_gradient(gphase_circuit(), 1, [([(0, "Z")], 1.0)], {"t": 0.6}, ["t"])
    assert g[0] == pytest.approx(-math.sin(0.6), abs=ATOL)
