"""Pins the C restatement (oracle/qarp_oracle.c) before it is trusted as the
checker: against the reference's own outputs (golden fixtures written from
the compiled reference, and the reference build itself where present),
against known-answer values from the reference's tests and the C++ standard,
and against closed forms / brute force that a plausible mistake would break.
CPU only."""
import math

import numpy as np
import pytest

import _fixtures as F
import _oracle as O
from openqarp_amd import workloads as W
from openqarp_amd._abi import pack_observable, pack_stream
from openqarp_amd.ir import Command, GateType, Param

REF = O.reference() is not None
needs_ref = pytest.mark.skipif(not REF, reason="reference build (oracle/_ref) not present")


# ── statevector / fusion ───────────────────────────────────────────────────

@pytest.mark.parametrize("name", ["hea12.npz", "hea16.npz", "hea8_unfused.npz"])
def test_hea_statevector_matches_reference_fixture(name):
    f = F.load(name)
    n = int(f["n"])
    w = W.c1_hea(n, int(f["layers"]), int(f["seed"]))
    k, kmin = (int(x) for x in f["fusion"])
    sv = O.statevector(pack_stream(w.commands), n, k, kmin)
    assert np.abs(sv - f["state"]).max() < 1e-12


@pytest.mark.parametrize("k", [2, 3, 4, 5])
def test_brickwork_fusion_widths_match_reference_fixture(k):
    f = F.load("brickwork10.npz")
    n = int(f["n"])
    cmds = W.brickwork_haar(n, int(f["layers"]), int(f["seed"]))
    sv = O.statevector(pack_stream(cmds), n, k, 0)
    assert np.abs(sv - f[f"state_k{k}"]).max() < 1e-12


def test_fused_stream_matches_reference_fixture():
    f = F.load("brickwork10.npz")
    n = int(f["n"])
    cmds = W.brickwork_haar(n, int(f["layers"]), int(f["seed"]))
    ours = O.fuse(pack_stream(cmds), n, 5, 0)
    assert len(ours) == len(f["fused_gates"])
    for i, (src, qs, m) in enumerate(ours):
        ref_q = tuple(int(q) for q in f["fused_qubits"][i] if q >= 0)
        if src >= 0:
            assert ref_q == cmds[src].qubits
        else:
            assert ref_q == qs
            d = m.shape[0]
            assert np.abs(m - f["fused_mats"][i][:d, :d]).max() < 1e-12


def _dense_unitary(cmds, n):
    """Brute force: the circuit's unitary by Kronecker embedding of each gate's
    closed-form matrix (conventions §2-§3), independent of the oracle's
    kernels and fusion."""
    U = np.eye(1 << n, dtype=np.complex128)
    for c in cmds:
        ps = pack_stream([c])
        k = len(c.qubits)
        g = np.zeros(1 << (2 * k), np.complex128)
        O.oracle().orc_gate_matrix(ps.ref, 0, g.ctypes.data, g.size)
        g = g.reshape(1 << k, 1 << k)
        E = np.zeros((1 << n, 1 << n), np.complex128)
        for col in range(1 << n):
            li = sum(((col >> q) & 1) << b for b, q in enumerate(c.qubits))
            rest = col & ~sum(1 << q for q in c.qubits)
            for lo in range(1 << k):
                row = rest | sum(((lo >> b) & 1) << q for b, q in enumerate(c.qubits))
                E[row, col] = g[lo, li]
        U = E @ U
    return U


def _closed_form(gate, params):
    c, s = math.cos(params[0] / 2), math.sin(params[0] / 2) if params else (1, 0)
    X = np.array([[0, 1], [1, 0]]); Y = np.array([[0, -1j], [1j, 0]]); Z = np.diag([1, -1])
    I = np.eye(2)
    if gate == GateType.Rx: return c * I - 1j * s * X
    if gate == GateType.Ry: return c * I - 1j * s * Y
    if gate == GateType.Rz: return c * I - 1j * s * Z
    if gate == GateType.RXX: return c * np.eye(4) - 1j * s * np.kron(X, X)
    if gate == GateType.RYY: return c * np.eye(4) - 1j * s * np.kron(Y, Y)
    if gate == GateType.RZZ: return c * np.eye(4) - 1j * s * np.kron(Z, Z)
    raise KeyError(gate)


@pytest.mark.parametrize("gate,nq", [(GateType.Rx, 1), (GateType.Ry, 1), (GateType.Rz, 1),
                                     (GateType.RXX, 2), (GateType.RYY, 2), (GateType.RZZ, 2)])
def test_rotation_matrices_are_exp_minus_i_theta_half_p(gate, nq):
    th = 0.8137
    c = Command(gate, qubits=list(range(nq)), params=[th])
    g = np.zeros(1 << (2 * nq), np.complex128)
    O.oracle().orc_gate_matrix(pack_stream([c]).ref, 0, g.ctypes.data, g.size)
    assert np.abs(g.reshape(1 << nq, 1 << nq) - _closed_form(gate, [th])).max() < 1e-15


def test_controlled_gate_convention_q0_control_q1_target():
    # CX(0, 1) on |q1 q0> = |01> (index 1) gives |11> (index 3).
    sv = O.statevector(pack_stream([Command(GateType.X, 0), Command(GateType.CX, 0, 1)]), 2, 0, 0)
    assert abs(sv[3] - 1) < 1e-15


@pytest.mark.parametrize("k", [0, 1, 2, 3, 4, 5])
def test_fused_statevector_equals_brute_force_unitary(k):
    rng = np.random.default_rng(5)
    n = 5
    gates = [GateType.H, GateType.Rx, GateType.Ry, GateType.Rz, GateType.CX, GateType.CZ, GateType.RZZ,
             GateType.SWAP, GateType.CRy, GateType.U, GateType.T, GateType.SX, GateType.iSWAP,
             GateType.ECR, GateType.CCX, GateType.CSWAP, GateType.CP, GateType.RYY]
    cmds = []
    for _ in range(60):
        g = gates[rng.integers(len(gates))]
        nq = 3 if g in (GateType.CCX, GateType.CSWAP) else 2 if g.value >= 16 else 1
        qs = [int(q) for q in rng.choice(n, nq, replace=False)]
        np_ = {GateType.U: 3}.get(g, 1 if g in (GateType.Rx, GateType.Ry, GateType.Rz, GateType.RZZ,
                                                  GateType.CRy, GateType.CP, GateType.RYY) else 0)
        cmds.append(Command(g, qubits=qs, params=list(rng.uniform(-3, 3, np_))))
    sv = O.statevector(pack_stream(cmds), n, k, 0)
    U = _dense_unitary(cmds, n)
    assert np.abs(sv - U[:, 0]).max() < 1e-12


@needs_ref
def test_random_circuits_match_reference_build():
    rng = np.random.default_rng(9)
    for trial in range(6):
        n = int(rng.integers(3, 11))
        cmds = W.hea(n, 3, int(rng.integers(1 << 30))).commands + W.brickwork_haar(n, 2, trial)
        ps = pack_stream(cmds)
        for k, kmin in ((0, 0), (1, 0), (3, 0), (5, 0), (3, 12)):
            a = O.statevector(ps, n, k, kmin)
            b = O.statevector(ps, n, k, kmin, lib="ref")
            assert np.abs(a - b).max() < 1e-12


# ── Pauli expectation ──────────────────────────────────────────────────────

def test_transition_matches_reference_fixture():
    f = F.load("transition8.npz")
    n = int(f["n"])
    po = pack_observable(F.observable(f))
    v = O.transition(f["bra"], f["ket"], n, po)
    assert abs(v - complex(f["value"])) < 1e-12


def test_transition_equals_dense_matrix_contraction():
    f = F.load("transition8.npz")
    n = int(f["n"])
    obs = F.observable(f)
    H = F.pauli_matrix(obs, n)
    expect = np.vdot(f["bra"], H @ f["ket"])
    assert abs(O.transition(f["bra"], f["ket"], n, pack_observable(obs)) - expect) < 1e-12


@pytest.mark.parametrize("n", [0, 1, 2, 5, 6, 7])
def test_transition_small_and_blocked_paths_agree_with_matrix(n):
    rng = np.random.default_rng(n)
    dim = 1 << n
    bra = rng.standard_normal(dim) + 1j * rng.standard_normal(dim)
    ket = rng.standard_normal(dim) + 1j * rng.standard_normal(dim)
    obs = W.random_pauli_sum(max(n, 1), 20, n, probs=(0.4, 0.2, 0.2, 0.2)) if n else [([], 0.7 - 0.2j)]
    obs = [([(q, l) for q, l in p if q < n], c) for p, c in obs]
    v = O.transition(bra, ket, n, pack_observable(obs))
    assert abs(v - np.vdot(bra, F.pauli_matrix(obs, n) @ ket)) < 1e-12


def test_h2_fci_energy_known_answer():
    """H2/STO-3G JW Hamiltonian of the reference's test fixture
    (tests/conftest.py:92-114): its ground-state expectation through the
    grouped transition is the FCI energy -1.1373060357534004 the reference
    asserts (tests/test_engines/test_expectation_kernel.py:538-546)."""
    obs = h2_jordan_wigner()
    H = F.pauli_matrix(obs, 4)
    w, v = np.linalg.eigh(H)
    assert w[0] == pytest.approx(-1.1373060357534004, abs=1e-10)
    e = O.transition(v[:, 0], v[:, 0], 4, pack_observable(obs))
    assert e.real == pytest.approx(-1.1373060357534004, abs=1e-12)


def h2_jordan_wigner():
    """JW image of the H2 fermion terms; a_p = Z_{<p} (X_p + iY_p)/2."""
    terms = [
        ((), 0.7199689944489797),
        (((0, 1), (0, 0)), -1.25633907300325), (((2, 1), (2, 0)), -0.47189600728114184),
        (((1, 1), (1, 0)), -1.25633907300325), (((3, 1), (3, 0)), -0.47189600728114184),
        (((2, 1), (0, 1), (0, 0), (2, 0)), 0.4836505304710652),
        (((3, 1), (1, 1), (1, 0), (3, 0)), 0.4836505304710652),
        (((1, 1), (0, 1), (0, 0), (1, 0)), 0.6757101548035163),
        (((2, 1), (1, 1), (1, 0), (2, 0)), 0.6645817302552967),
        (((1, 1), (0, 1), (2, 0), (3, 0)), 0.18093119978423133),
        (((2, 1), (1, 1), (0, 0), (3, 0)), -0.1809311997842314),
        (((3, 1), (0, 1), (1, 0), (2, 0)), -0.1809311997842314),
        (((3, 1), (2, 1), (0, 0), (1, 0)), 0.18093119978423136),
        (((3, 1), (0, 1), (0, 0), (3, 0)), 0.6645817302552967),
        (((3, 1), (2, 1), (2, 0), (3, 0)), 0.6985737227320175),
    ]
    # Pauli algebra on dicts {tuple(letters): coeff}
    mul = {("I", "I"): (1, "I"), ("I", "X"): (1, "X"), ("I", "Y"): (1, "Y"), ("I", "Z"): (1, "Z"),
           ("X", "I"): (1, "X"), ("Y", "I"): (1, "Y"), ("Z", "I"): (1, "Z"),
           ("X", "X"): (1, "I"), ("Y", "Y"): (1, "I"), ("Z", "Z"): (1, "I"),
           ("X", "Y"): (1j, "Z"), ("Y", "X"): (-1j, "Z"), ("Y", "Z"): (1j, "X"),
           ("Z", "Y"): (-1j, "X"), ("Z", "X"): (1j, "Y"), ("X", "Z"): (-1j, "Y")}

    def ladder(p, dag):
        out = {}
        for let, c in (("X", 0.5), ("Y", 0.5j if not dag else -0.5j)):
            s = ["Z"] * p + [let] + ["I"] * (3 - p)
            out[tuple(s)] = c
        return out

    def product(a, b):
        out = {}
        for sa, ca in a.items():
            for sb, cb in b.items():
                ph, s = 1, []
                for x, y in zip(sa, sb):
                    f, r = mul[(x, y)]
                    ph *= f
                    s.append(r)
                out[tuple(s)] = out.get(tuple(s), 0) + ca * cb * ph
        return out

    total = {}
    for ops, coeff in terms:
        acc = {("I",) * 4: coeff}
        for p, d in ops:
            acc = product(acc, ladder(p, d == 1))
        for s, c in acc.items():
            total[s] = total.get(s, 0) + c
    obs = []
    for s, c in total.items():
        if abs(c) > 1e-14:
            obs.append(([(q, l) for q, l in enumerate(s) if l != "I"], complex(c)))
    return obs


def test_analytic_expectations_product_ghz_and_plus_i():
    # |+>^6: <X_q> = 1, <Z_q> = <Y_q> = 0 (test_expectation_kernel.py:176-190)
    n = 6
    plus = pack_stream([Command(GateType.H, q) for q in range(n)])
    sv = O.statevector(plus, n, 0, 0)
    for q in range(n):
        assert O.transition(sv, sv, n, pack_observable([([(q, "X")], 1.0)])).real == pytest.approx(1, abs=1e-12)
        assert O.transition(sv, sv, n, pack_observable([([(q, "Y")], 1.0)])).real == pytest.approx(0, abs=1e-12)
    # GHZ_7: <X...X> = 1, <Y0 Y1 X...> = -1 (:193-205)
    n = 7
    ghz = [Command(GateType.H, 0)] + [Command(GateType.CX, 0, q) for q in range(1, n)]
    sv = O.statevector(pack_stream(ghz), n, 0, 0)
    xall = [([(q, "X") for q in range(n)], 1.0)]
    yyx = [([(0, "Y"), (1, "Y")] + [(q, "X") for q in range(2, n)], 1.0)]
    assert O.transition(sv, sv, n, pack_observable(xall)).real == pytest.approx(1, abs=1e-12)
    assert O.transition(sv, sv, n, pack_observable(yyx)).real == pytest.approx(-1, abs=1e-12)
    # |+i> = S H |0>: <Y> = +1 (:208-216)
    sv = O.statevector(pack_stream([Command(GateType.H, 0), Command(GateType.S, 0)]), 1, 0, 0)
    assert O.transition(sv, sv, 1, pack_observable([([(0, "Y")], 1.0)])).real == pytest.approx(1, abs=1e-12)


def test_transition_rejects_bad_input():
    sv = np.zeros(4, np.complex128); sv[0] = 1
    for obs in ([([(2, "X")], 1.0)], [([(0, "X"), (0, "Z")], 1.0)], [([(0, "Q")], 1.0)]):
        with pytest.raises(O.OracleError):
            O.transition(sv, sv, 2, pack_observable(obs))


def test_apply_observable_equals_matrix_vector():
    f = F.load("transition8.npz")
    n = int(f["n"])
    obs = F.observable(f)
    out = O.apply_observable(f["ket"], n, pack_observable(obs))
    assert np.abs(out - F.pauli_matrix(obs, n) @ f["ket"]).max() < 1e-12


# ── sampling ───────────────────────────────────────────────────────────────

def test_mt19937_known_answer():
    """C++ [rand.predef]: the 10000th output of a default-constructed
    std::mt19937 (seed 5489) is 4123659995."""
    assert int(O.mt19937(5489, 10000)[-1]) == 4123659995


def test_mt19937_and_uniforms_match_reference_fixture():
    f = F.load("mt19937.npz")
    assert (O.mt19937(5489, 2000) == f["words_5489"]).all()
    assert (O.mt19937(1234, 2000) == f["words_1234"]).all()
    assert (O.uniforms(1234, 2000) == f["uniforms_1234"]).all()


def test_uniform_is_two_words_over_2_pow_64():
    w = O.mt19937(42, 20).astype(np.float64)
    u = O.uniforms(42, 10)
    assert (u == (w[0::2] + w[1::2] * 4294967296.0) / 18446744073709551616.0).all()


@pytest.mark.parametrize("name", ["sample10.npz", "hea12.npz", "hea16.npz"])
def test_samples_match_reference_fixture(name):
    f = F.load(name)
    n = int(f["n"])
    out = O.sample(f["state"], n, int(f["shots"]), int(f["shot_seed"]))
    assert (out == f["samples"]).all()


def test_cdf_is_the_sequential_recurrence():
    f = F.load("sample10.npz")
    sv = f["state"]
    p = sv.real * sv.real + sv.imag * sv.imag
    s = 0.0
    for v in p:
        s = s + float(v)
    run, cdf = 0.0, []
    for v in p:
        run = run + float(v) / s
        cdf.append(run)
    cdf[-1] = 1.0
    assert (O.cdf(sv, int(f["n"])) == np.array(cdf)).all()


def test_sampling_distribution_is_right():
    rng = np.random.default_rng(3)
    amp = rng.standard_normal(8) + 1j * rng.standard_normal(8)
    amp /= np.linalg.norm(amp)
    out = O.sample(amp, 3, 200000, 5)
    freq = np.bincount(out.astype(np.int64), minlength=8) / 200000
    assert np.abs(freq - np.abs(amp) ** 2).max() < 0.005


# ── adjoint gradient ───────────────────────────────────────────────────────

def _fd_energy(w, n, eps=1e-6):
    vals = np.array([w.values[s] for s in w.symbols])
    po = pack_observable(w.observable)

    def energy(v):
        cmds = [c.substitute(dict(zip(w.symbols, v))) for c in w.commands]
        sv = O.statevector(pack_stream(cmds), n, 0, 0)
        return O.transition(sv, sv, n, po).real

    g = np.zeros(len(vals))
    for i in range(len(vals)):
        vp, vm = vals.copy(), vals.copy()
        vp[i] += eps
        vm[i] -= eps
        g[i] = (energy(vp) - energy(vm)) / (2 * eps)
    return g


def test_gradient_matches_reference_fixture_and_finite_differences():
    f = F.load("gradient8.npz")
    n = int(f["n"])
    w = W.c3_gradient(n, int(f["layers"]), int(f["seed"]), int(f["n_terms"]))
    idx = {s: i for i, s in enumerate(w.symbols)}
    g = O.run_gradient(pack_stream(w.commands, w.values, idx), n, pack_observable(w.observable), len(idx))
    assert np.abs(g - f["grad"]).max() < 1e-12
    assert np.abs(g - _fd_energy(w, n)).max() < 1e-7


def test_gradient_parameter_shift_closed_form():
    """RY/RZ occurrences: dE/dx = [E(x + pi/2) - E(x - pi/2)] / 2
    (_gradients.py:258-291, two-term rule with coefficient 1/2)."""
    w = W.c3_gradient(5, 2, 8, 20)
    n = 5
    idx = {s: i for i, s in enumerate(w.symbols)}
    po = pack_observable(w.observable)
    g = O.run_gradient(pack_stream(w.commands, w.values, idx), n, po, len(idx))

    def energy(vals):
        sv = O.statevector(pack_stream([c.substitute(vals) for c in w.commands]), n, 0, 0)
        return O.transition(sv, sv, n, po).real

    for s in w.symbols[:10]:
        vp = dict(w.values); vp[s] += math.pi / 2
        vm = dict(w.values); vm[s] -= math.pi / 2
        assert g[idx[s]] == pytest.approx(0.5 * (energy(vp) - energy(vm)), abs=1e-12)


def gphase_circuit():
    """RY(t) then GPhase(t) on one qubit, <Z> = cos t: the global phase drops
    out, so dE/dt = -sin t (src/qarpx/tests/test_gradient_adjoint.cpp:168-179)."""
    return [Command(GateType.Ry, 0, Param.symbol("t")), Command(GateType.GPhase, params=[Param.symbol("t")])]


def test_gphase_gradient_known_answer():
    ps = pack_stream(gphase_circuit(), {"t": 0.6}, {"t": 0})
    g = O.run_gradient(ps, 1, pack_observable([([(0, "Z")], 1.0)]), 1)
    assert g[0] == pytest.approx(-math.sin(0.6), abs=1e-12)


def test_gradient_through_every_generator_rule_matches_fd():
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
    rng = np.random.default_rng(4)
    vals = dict(zip(names, rng.uniform(-2, 2, 12)))
    obs = [([(0, "X"), (1, "Y")], 0.7), ([(2, "Z")], -0.4), ([(0, "Y"), (2, "X")], 0.3)]
    w = W.Workload(n, cmds, obs, vals, names)
    g = O.run_gradient(pack_stream(cmds, vals, {s: i for i, s in enumerate(names)}), n,
                       pack_observable(obs), len(names))
    assert np.abs(g - _fd_energy(w, n)).max() < 1e-7
    if REF:
        gr = O.run_gradient(pack_stream(cmds, vals, {s: i for i, s in enumerate(names)}), n,
                            pack_observable(obs), len(names), values=np.array([vals[s] for s in names]),
                            lib="ref")
        assert np.abs(g - gr).max() < 1e-12


def test_qaoa_fixture_state_energy_and_gradient():
    f = F.load("qaoa10.npz")
    n, p, seed = int(f["n"]), int(f["p"]), int(f["seed"])
    w = W.qaoa_maxcut(n, p, seed, symbolic=True)
    idx = {s: i for i, s in enumerate(w.symbols)}
    ps = pack_stream(w.commands, w.values, idx)
    po = pack_observable(w.observable)
    sv = O.statevector(ps, n, 3, 12)
    assert np.abs(sv - f["state"]).max() < 1e-12
    assert O.transition(sv, sv, n, po).real == pytest.approx(float(f["energy"]), abs=1e-12)
    assert np.abs(O.run_gradient(ps, n, po, len(idx)) - f["grad"]).max() < 1e-12


def _generator_circuit():
    """Every differentiable gate type once, with affine and compound angles."""
    t = [Param.symbol(f"t{i}") for i in range(12)]
    cmds = [Command(GateType.H, 0), Command(GateType.H, 1), Command(GateType.H, 2),
            Command(GateType.Rx, 0, t[0]), Command(GateType.Ry, 1, t[1] * 2.0 + 0.3),
            Command(GateType.Rz, 2, t[2]), Command(GateType.P, 0, t[3]),
            Command(GateType.CRx, 0, 1, t[4]), Command(GateType.CRy, 1, 2, t[5]),
            Command(GateType.CRz, 2, 0, t[6]), Command(GateType.CP, 0, 2, t[7]),
            Command(GateType.RXX, 0, 1, t[8]), Command(GateType.RYY, 1, 2, t[9]),
            Command(GateType.RZZ, 0, 2, t[10] - t[0]), Command(GateType.GPhase, params=[t[11]]),
            Command(GateType.S, 1), Command(GateType.SX, 2), Command(GateType.iSWAP, 0, 1)]
    return cmds, [f"t{i}" for i in range(12)]


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_gradient_phi_is_the_derivative_of_2re_phi_psi(seed):
    """run_gradient_phi with a fixed |phi> returns d/dtheta 2 Re<phi|psi(theta)>
    (run_gradient_core with the caller's phi, qarp_simulator.cpp:1626-1639):
    central differences of that closed form, from |0> and from an initial
    state; the same function equals the reference build where present."""
    n = 3
    cmds, names = _generator_circuit()
    rng = np.random.default_rng(seed)
    vals = dict(zip(names, rng.uniform(-2, 2, len(names))))
    index = {s: i for i, s in enumerate(names)}
    phi = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    init = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    init /= np.linalg.norm(init)
    for start in (None, init):
        g = O.run_gradient_phi(pack_stream(cmds, vals, index), n, phi, len(names), init=start)

        def f(v):
            sv = O.statevector(pack_stream([c.substitute(v) for c in cmds]), n, 0, 0, init=start)
            return 2.0 * np.vdot(phi, sv).real

        h = 1e-6
        fd = []
        for s in names:
            vp = dict(vals); vp[s] += h
            vm = dict(vals); vm[s] -= h
            fd.append((f(vp) - f(vm)) / (2 * h))
        assert np.abs(g - np.array(fd)).max() < 1e-7
        if REF:
            gr = O.run_gradient_phi(pack_stream(cmds, vals, index), n, phi, len(names),
                                    values=np.array([vals[s] for s in names]), init=start, lib="ref")
            assert np.abs(g - gr).max() < 1e-12


@needs_ref
def test_gradient_phi_matches_reference_build_on_random_ansatz():
    """A 9-qubit, 2-layer RY/RZ + CZ ansatz with phi = <bra|psi> |bra> (the
    OVERLAP adjoint's phi, _gradients.py:522-544): oracle equals the reference."""
    n = 9
    w = W.c3_ansatz(n, 2, 17)
    index = {s: i for i, s in enumerate(w.symbols)}
    ps = pack_stream(w.commands, w.values, index)
    rng = np.random.default_rng(5)
    bra = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    bra /= np.linalg.norm(bra)
    psi = O.statevector(pack_stream([c.substitute(w.values) for c in w.commands]), n, 0, 0)
    phi = np.vdot(bra, psi) * bra
    vals = np.array([w.values[s] for s in w.symbols])
    g = O.run_gradient_phi(ps, n, phi, len(vals))
    gr = O.run_gradient_phi(ps, n, phi, len(vals), values=vals, lib="ref")
    assert np.abs(g - gr).max() < 1e-12


def test_c1_full_size_fixture_against_oracle():
    """The 20-qubit HEA fixture (reference build, full size): the oracle's
    state at the fingerprint indices, <H>, <Z_q> and the 4000 shots."""
    f = F.load("c1_hea20.npz")
    n = int(f["n"])
    w = W.c1_hea(n, int(f["layers"]), int(f["seed"]))
    sv = O.statevector(pack_stream(w.commands), n, 3, 12)
    assert np.abs(sv[f["idx"].astype(np.int64)] - f["amps"]).max() < 1e-12
    p = np.abs(sv) ** 2
    bits = np.arange(1 << n)
    zq = np.array([p[(bits >> q) & 1 == 0].sum() - p[(bits >> q) & 1 == 1].sum() for q in range(n)])
    assert np.abs(zq - f["zq"]).max() < 1e-12
    assert O.transition(sv, sv, n, pack_observable(w.observable)).real == pytest.approx(float(f["energy"]),
                                                                                         abs=1e-12)
    assert np.array_equal(O.sample(sv, n, int(f["shots"]), int(f["shot_seed"])), f["samples"])
