"""The five workloads at their stated sizes against the reference build.

The fixtures (tests/golden/c*_*.npz, written by tests/golden/make_golden.py
--full from oracle/_ref) hold fingerprints of states too large to commit:
the amplitudes at 4096 seeded indices, the norm, <Z_q> for every qubit and
the workload's observable.  Bars: 1e-12 absolute for complex128 amplitudes
and expectation values, bit-exact shots, 1e-8 for adjoint gradients (against
the reference adjoint and against central finite differences); complex64
states report their error against the complex128 reference.
"""
import numpy as np
import pytest

import _fixtures as F
from openqarp_amd import workloads as W
from openqarp_amd.simulator import QarpSimulator

pytestmark = pytest.mark.gpu

ATOL = 1e-12
GRAD_TOL = 1e-8


def z_expectations(sim, st, n):
    return np.array([sim.expectation_state(st, [([(q, "Z")], 1.0)]) for q in range(n)])


def check_fingerprint(sim, st, f, atol=ATOL, ztol=ATOL):
    """Amplitudes at the fixture's indices within atol; the norm and every
    <Z_q> within ztol."""
    n = st.n_qubits
    amps = st.download()[f["idx"].astype(np.int64)]
    assert np.abs(amps - f["amps"]).max() < atol
    assert abs(st.norm2() - float(f["norm"])) < ztol
    assert np.abs(z_expectations(sim, st, n) - f["zq"]).max() < ztol


def test_c1_hea20_state_energy_and_shots():
    """20-qubit HEA, 10 layers (C1): amplitudes, <H>, 4000 shots bit-exact."""
    f = F.load("c1_hea20.npz")
    n = int(f["n"])
    w = W.c1_hea(n, int(f["layers"]), int(f["seed"]))
    sim = QarpSimulator()
    st = sim.new_state(n)
    sim.apply(w.commands, st)
    check_fingerprint(sim, st, f)
    assert abs(sim.expectation_state(st, w.observable) - float(f["energy"])) < ATOL
    shots = sim.sample_state(st, int(f["shots"]), int(f["shot_seed"]))
    assert np.array_equal(shots, f["samples"])


@pytest.mark.parametrize("fusion,tile,low", [(4, 12, 3), (5, 11, 3), (3, 12, 4)])
def test_c2_brickwork30_fingerprint_and_energy(fusion, tile, low):
    """30-qubit Haar brickwork (C2, 16 GiB): the tuned headline shape (fusion
    4, 2^12-amplitude tiles, relabeled passes, FP64 matrix cores) and two
    others against the reference's fusion-5 state; the 2000-string <H> once."""
    f = F.load("c2_brickwork30.npz")
    n = int(f["n"])
    w = W.c2_brickwork(n, int(f["layers"]), int(f["seed"]), int(f["n_terms"]))
    sim = QarpSimulator(tile_qubits=tile, low_qubits=low)
    sim.fusion_max_qubits, sim.fusion_min_qubits = fusion, 12
    st = sim.new_state(n)
    stats = sim.apply(w.commands, st)
    assert stats["passes"] > 0
    check_fingerprint(sim, st, f)
    if fusion == 4:
        assert abs(sim.expectation_state(st, w.observable) - float(f["energy"])) < ATOL


def test_c3_gradient24_against_reference_and_fd():
    """24-qubit, 384-parameter ansatz with 1000 local strings (C3): energy and
    the full adjoint gradient against the reference adjoint, eight seeded
    components against central finite differences (h = 1e-6)."""
    f = F.load("c3_gradient24.npz")
    n = int(f["n"])
    w = W.c3_gradient(n, int(f["layers"]), int(f["seed"]), int(f["n_terms"]))
    assert np.array_equal(np.array([w.values[s] for s in w.symbols]), f["values"])
    sim = QarpSimulator()
    grad, energy = sim.run_gradient_energy(w.commands, n, w.observable, w.values, w.symbols)
    assert abs(energy - float(f["energy"])) < ATOL
    assert np.abs(grad - f["grad"]).max() < GRAD_TOL
    st = sim.new_state(n)
    sim.apply([c.substitute(w.values) for c in w.commands], st)
    check_fingerprint(sim, st, f)
    rng = np.random.default_rng(11)
    h = 1e-6
    for k in rng.choice(len(w.symbols), 8, replace=False):
        s = w.symbols[int(k)]
        e = []
        for sign in (1.0, -1.0):
            vals = dict(w.values)
            vals[s] += sign * h
            e.append(sim.expectation([c.substitute(vals) for c in w.commands], n, w.observable))
        assert abs((e[0] - e[1]) / (2 * h) - grad[int(k)]) < GRAD_TOL


def test_c4_batch512_subset_against_reference():
    """The 512-instance batch (C4, 12..22 qubits, depth 4..16) in one batched
    sweep: the seeded subset (including the widest instances) against the
    reference's energies and gradients; complex64 within float32 accuracy."""
    f = F.load("c4_batch_subset.npz")
    ws = W.batched_vqe(int(f["n_instances"]), seed=int(f["seed"]))
    inst = [(w.commands, w.n_qubits, w.observable, w.values, w.symbols) for w in ws]
    e128, g128 = QarpSimulator().batch_energy_grad(inst)
    offs = f["grad_offsets"]
    for j, i in enumerate(f["pick"]):
        i = int(i)
        assert ws[i].n_qubits == int(f["widths"][j])
        assert abs(e128[i] - float(f["energies"][j])) < ATOL
        assert np.abs(g128[i] - f["grad"][offs[j]:offs[j + 1]]).max() < GRAD_TOL
    e64, g64 = QarpSimulator(precision="complex64").batch_energy_grad(inst)
    for i in range(len(inst)):
        scale = sum(abs(complex(c)) for _, c in inst[i][2])
        assert abs(e64[i] - e128[i]) < 2e-5 * scale
        assert np.abs(g64[i] - g128[i]).max() < 2e-5 * scale


@pytest.mark.parametrize("precision", ["complex128", "complex64"])
def test_c5_qaoa24_against_reference(precision):
    """QAOA MaxCut, p = 4, on the seeded 3-regular graph construction at 24
    qubits (C5's family at a size the reference can run): <C>, its adjoint
    gradient and the amplitudes; complex64 against the same complex128
    reference at float32 accuracy."""
    f = F.load("c5_qaoa24.npz")
    n = int(f["n"])
    w = W.qaoa_maxcut(n, int(f["p"]), int(f["seed"]), symbolic=True)
    sim = QarpSimulator(precision=precision)
    grad, energy = sim.run_gradient_energy(w.commands, n, w.observable, w.values, w.symbols)
    st = sim.new_state(n)
    sim.apply([c.substitute(w.values) for c in w.commands], st)
    if precision == "complex128":
        assert abs(energy - float(f["energy"])) < ATOL
        assert np.abs(grad - f["grad"]).max() < GRAD_TOL
        check_fingerprint(sim, st, f)
    else:
        assert abs(energy - float(f["energy"])) < 1e-4
        assert np.abs(grad - f["grad"]).max() < 1e-4
        check_fingerprint(sim, st, f, atol=1e-6, ztol=1e-4)
