"""Structured programs (qx.Program kernels run by QarpSimulator::
program_statevector / program_run, src/qarpx/src/simulator/program.cpp and
qarp_simulator.cpp:789-888).

CPU: the oracle's permutation and controlled-power kernels are pinned by
the reference build and its golden fixture, and independently by the dense
apply of the same operators (a permutation matrix on the subset; diag(I,
U^e) on targets + control); the Program class keeps the reference's
validation, folding and bookkeeping.  GPU (marked): the HIP kernels against
the oracle — permutations bit-exact, amplitudes to 1e-12 (complex128)."""
from pathlib import Path

import numpy as np
import pytest

import _oracle as O
from openqarp_amd import workloads as W
from openqarp_amd._abi import pack_stream
from openqarp_amd.ir import Command, GateType, Param
from openqarp_amd.program import Program
from test_gpu_parity import random_circuit

GOLDEN = Path(__file__).resolve().parent / "golden"
ATOL = 1e-12


def haar(d, rng):
    return W.haar_unitary(d, rng)


def random_state(n, rng):
    v = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    return v / np.linalg.norm(v)


def dense_apply(psi, n, qubits, m):
    """The operator as one Custom op through the oracle's dense apply."""
    return O.statevector(pack_stream([Command(GateType.Custom, qubits=list(qubits), unitary=m)]), n, 0, 0,
                         init=psi)


def perm_matrix(table):
    d = len(table)
    P = np.zeros((d, d), dtype=np.complex128)
    P[np.asarray(table), np.arange(d)] = 1.0
    return P


# ── the oracle, pinned ──────────────────────────────────────────────────────

def test_oracle_program_matches_reference_fixture():
    d = np.load(GOLDEN / "program10.npz")
    prog = W.structured_program(int(d["n"]), int(d["seed"]))
    assert prog.kinds() == list(d["kinds"])
    got = O.program_statevector(prog.kernels(), int(d["n"]))
    assert np.abs(got - d["state"]).max() < ATOL


@pytest.mark.parametrize("n,qubits", [(6, [0, 1, 2, 3, 4, 5]), (7, [3]), (7, [6, 0, 4]), (8, [2, 7, 5, 1]),
                                      (5, [4, 3, 2, 1, 0])])
def test_oracle_permutation_is_the_permutation_operator(n, qubits):
    rng = np.random.default_rng(n * 10 + len(qubits))
    table = rng.permutation(1 << len(qubits))
    psi = random_state(n, rng)
    got = O.permute(psi, n, qubits, table)
    assert np.abs(got - dense_apply(psi, n, qubits, perm_matrix(table))).max() < 1e-15
    # Inverse table undoes it exactly; identity table changes nothing.
    inv = np.argsort(table)
    assert np.array_equal(O.permute(got, n, qubits, inv), psi)
    assert np.array_equal(O.permute(psi, n, qubits, np.arange(1 << len(qubits))), psi)


@pytest.mark.parametrize("exps", [[1, 2, 4], [0, 3, 6], [5, 0, 1], [13, 7, 2]])
def test_oracle_controlled_powers_are_controlled_matrix_powers(exps):
    n = 7
    rng = np.random.default_rng(sum(exps))
    U = haar(4, rng)
    controls, targets = [6, 2, 4], [5, 0]
    psi = random_state(n, rng)
    got = O.controlled_powers(psi, n, controls, exps, targets, U)
    want = psi
    for c, e in zip(controls, exps):
        blk = np.eye(8, dtype=np.complex128)
        blk[4:, 4:] = np.linalg.matrix_power(U, e)
        want = dense_apply(want, n, targets + [c], blk)
    assert np.abs(got - want).max() < 1e-13


@pytest.mark.skipif(O.reference() is None, reason="reference build absent")
@pytest.mark.parametrize("seed", range(4))
def test_oracle_program_matches_reference_build(seed):
    rng = np.random.default_rng(100 + seed)
    n = 9
    p = Program()
    p.add_gates(random_circuit(n, 25, seed))
    for _ in range(3):   # consecutive permutations fold
        qs = [int(q) for q in rng.choice(n, size=int(rng.integers(1, n + 1)), replace=False)]
        p.add_permutation(qs, rng.permutation(1 << len(qs)))
    p.add_dense([int(q) for q in rng.choice(n, size=4, replace=False)], haar(16, rng))
    cq = [int(q) for q in rng.choice(n, size=6, replace=False)]
    p.add_controlled_powers(cq[:3], [int(e) for e in rng.integers(0, 40, size=3)], cq[3:], haar(8, rng))
    p.add_permutation(list(range(n)), rng.permutation(1 << n))
    init = random_state(n, rng)
    a = O.program_statevector(p.kernels(), n, init=init)
    b = O.program_statevector(p.kernels(), n, init=init, lib="ref")
    assert np.abs(a - b).max() < ATOL


# ── Program bookkeeping and validation (program.cpp:486-605) ─────────────────

def test_program_kernels_fold_and_report():
    p = Program()
    p.add_gates([Command(GateType.H, 0), Command(GateType.Measure, 1, cbits=[0]), Command(GateType.Barrier, 0)])
    p.add_permutation([0, 1], [1, 0, 3, 2])
    p.add_permutation([2], [1, 0])                 # folds: one kernel over qubits 0, 1, 2
    assert p.kinds() == ["gates", "permutation"] and len(p) == 2
    _, qs, table = p.kernels()[1]
    assert qs == [0, 1, 2]
    assert list(table) == [5, 4, 7, 6, 1, 0, 3, 2]
    p.add_dense([3], np.eye(2))
    p.add_permutation([0], [1, 0])                 # not after a permutation: a kernel of its own
    p.add_controlled_powers([5], [3], [4], np.eye(2))
    assert p.kinds() == ["gates", "permutation", "dense", "permutation", "controlled_powers"]
    assert p.min_register_width == 6
    assert Program().min_register_width == 0
    q = p.without_measurements()
    assert [c.gate for c in q.kernels()[0][1]] == [GateType.H]
    assert len(p.kernels()[0][1]) == 3


def test_program_fold_stops_at_the_table_cap():
    p = Program()
    p.add_permutation(list(range(20)), np.arange(1 << 20)[::-1])
    p.add_permutation(list(range(20, 27)), np.arange(1 << 7))
    assert p.kinds() == ["permutation", "permutation"]


def test_program_substituted_binds_gate_parameters():
    p = Program()
    p.add_gates([Command(GateType.Rx, 0, params=[Param.symbol("a")])])
    p.add_permutation([0], [1, 0])
    q = p.substituted({"a": 0.5})
    assert float(q.kernels()[0][1][0].params[0].value()) == 0.5
    assert q.kinds() == p.kinds()


@pytest.mark.parametrize("call,msg", [
    (lambda p: p.add_permutation([0, 0], [0, 1, 2, 3]), "add_permutation: qubits must be distinct"),
    (lambda p: p.add_permutation([], []), "add_permutation: 1..26 qubits required"),
    (lambda p: p.add_permutation([0], [0, 1, 2]), "add_permutation: table length 3 for 1 qubit(s)"),
    (lambda p: p.add_permutation([0, 1], [0, 1, 1, 2]), "add_permutation: table is not a bijection"),
    (lambda p: p.add_permutation([0], [0, 2]), "add_permutation: table is not a bijection"),
    (lambda p: p.add_dense([], np.eye(1)), "add_dense: 1..8 qubits required"),
    (lambda p: p.add_dense([0, 1], np.eye(2)), "add_dense: matrix is not 2^k square"),
    (lambda p: p.add_controlled_powers([], [], [0], np.eye(2)), "add_controlled_powers: controls and targets required"),
    (lambda p: p.add_controlled_powers([1], [1, 2], [0], np.eye(2)), "add_controlled_powers: one exponent per control"),
    (lambda p: p.add_controlled_powers([0], [1], [0], np.eye(2)), "add_controlled_powers: qubits must be distinct"),
    (lambda p: p.add_controlled_powers([1], [1], [0], np.eye(4)), "add_controlled_powers: matrix is not 2^|targets| square"),
])
def test_program_validation_messages(call, msg):
    with pytest.raises(ValueError) as e:
        call(Program())
    assert str(e.value) == msg


# ── the HIP path ────────────────────────────────────────────────────────────

def _sim(**kw):
    from openqarp_amd.simulator import QarpSimulator
    return QarpSimulator(**kw)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["complex128", "complex64"])
def test_program_statevector_matches_reference_fixture(precision):
    d = np.load(GOLDEN / "program10.npz")
    prog = W.structured_program(int(d["n"]), int(d["seed"]))
    got = _sim(precision=precision).program_statevector(prog, int(d["n"]))
    assert np.abs(got - d["state"]).max() < (ATOL if precision == "complex128" else 2e-6)


@pytest.mark.gpu
@pytest.mark.parametrize("seed", range(3))
def test_program_kernels_against_oracle(seed):
    """Every kernel kind at 14 qubits, tiles of 2^9 amplitudes (several
    passes): permutations over subsets and the whole register, dense blocks
    of 1..8 qubits, controlled powers on 1..7 targets (tile-pass ops)."""
    rng = np.random.default_rng(200 + seed)
    n = 14
    p = Program()
    p.add_gates(random_circuit(n, 40, seed))
    for k in (1, 5, n):
        qs = [int(q) for q in rng.choice(n, size=k, replace=False)]
        p.add_permutation(qs, rng.permutation(1 << k))
    for k in (1, 3, 6, 8):
        p.add_dense([int(q) for q in rng.choice(n, size=k, replace=False)], haar(1 << k, rng))
    for m in (1, 2, 5, 7):
        qs = [int(q) for q in rng.choice(n, size=m + 3, replace=False)]
        p.add_controlled_powers(qs[m:], [int(e) for e in rng.integers(0, 1 << 12, size=3)], qs[:m],
                                haar(1 << m, rng))
    p.add_gates(random_circuit(n, 30, seed + 10))
    init = random_state(n, rng)
    want = O.program_statevector(p.kernels(), n, init=init)
    got = _sim(tile_qubits=9).program_statevector(p, n, initial_state=init)
    assert np.abs(got - want).max() < ATOL


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["complex128", "complex64"])
def test_permutation_is_bit_exact_at_22_qubits(precision):
    """A permutation moves amplitudes without arithmetic: the device result
    equals numpy's gather of the same table bit for bit (20-qubit table, and
    a whole-register table at 16 qubits)."""
    from openqarp_amd.simulator import DeviceState
    rng = np.random.default_rng(5)
    dt = np.complex128 if precision == "complex128" else np.complex64
    for n, qs in ((22, [int(q) for q in rng.choice(22, size=20, replace=False)]), (16, list(range(16)))):
        table = rng.permutation(1 << len(qs)).astype(np.uint64)
        psi = random_state(n, rng).astype(dt)
        sim = _sim(precision=precision)
        st = DeviceState(n, precision).upload(psi)
        sim.execute_program(_perm_program(qs, table), st)
        got = st.download().astype(dt)
        want = np.empty_like(psi)
        want[_perm_targets(n, qs, table)] = psi
        assert np.array_equal(got, want)


def _perm_program(qs, table):
    p = Program()
    p.add_permutation(qs, table)
    return p


def _perm_targets(n, qs, table):
    i = np.arange(1 << n, dtype=np.uint64)
    x = np.zeros_like(i)
    for b, q in enumerate(qs):
        x |= ((i >> np.uint64(q)) & np.uint64(1)) << np.uint64(b)
    y = table[x]
    keep = np.uint64(((1 << 64) - 1) ^ sum(1 << q for q in qs))
    j = i & keep
    for b, q in enumerate(qs):
        j |= ((y >> np.uint64(b)) & np.uint64(1)) << np.uint64(q)
    return j


@pytest.mark.gpu
def test_program_run_samples_bit_exact():
    """program_run: the program's state sampled with the run() seed stream;
    outcomes equal the oracle's sampling of the oracle's state."""
    n = 10
    prog = W.structured_program(n, 8)
    meas = [Command(GateType.Measure, q, cbits=[q]) for q in range(n)]
    prog.add_gates(meas)
    res = _sim().program_run(prog, n, 3000, seed=1234)
    sv = O.program_statevector(W.structured_program(n, 8).kernels(), n)
    outcomes = O.sample(sv, n, 3000, 1234)
    vals, cnt = np.unique(outcomes, return_counts=True)
    assert res.counts == {int(v): int(c) for v, c in zip(vals, cnt)}
    assert len(res.cbit_history) == 3000


@pytest.mark.gpu
def test_program_errors():
    from openqarp_amd.errors import CapabilityError
    sim = _sim()
    p = Program()
    p.add_permutation([5], [1, 0])
    with pytest.raises(ValueError, match="program_statevector: a kernel acts on qubit 5 of a 4-qubit register"):
        sim.program_statevector(p, 4)
    q = Program()
    q.add_gates([Command(GateType.Measure, 0, cbits=[0]), Command(GateType.X, 0)])
    with pytest.raises(RuntimeError, match="statevector is not deterministic"):
        sim.program_statevector(q, 2)
    with pytest.raises(ValueError, match="measurements are not terminal"):
        sim.program_run(q, 2, 10)
    q.add_permutation([1], [1, 0])
    with pytest.raises(ValueError, match="outside the program's last gates kernel"):
        sim.program_run(q, 2, 10)
    # 15 targets: rejected by the library before the matrix is read.
    import ctypes as C
    from openqarp_amd import _lib
    from openqarp_amd.simulator import DeviceState
    st = DeviceState(16, "complex128")
    cs, es, ts = np.array([15], np.uint32), np.array([1], np.uint64), np.arange(15, dtype=np.uint32)
    cfg = sim._config()
    with pytest.raises(CapabilityError, match="at most 14 targets"):
        _lib.check(_lib.load().qhip_state_controlled_powers(C.byref(cfg), st.handle, cs.ctypes.data, es.ctypes.data,
                                                            1, ts.ctypes.data, 15, np.zeros(2).ctypes.data))


@pytest.mark.gpu
@pytest.mark.parametrize("m,n,precision,exps,low", [
    (8, 12, "complex128", [5, 0, 6, 1], False), (9, 12, "complex64", [3, 1], False),
    (10, 13, "complex128", [1], False), (11, 13, "complex64", [1], False), (12, 13, "complex128", [1], False),
    (8, 18, "complex128", [3, 1], True), (8, 18, "complex128", [6], False), (9, 17, "complex64", [2], True),
    (13, 15, "complex128", [1], True)])
def test_wide_controlled_powers_against_oracle(m, n, precision, exps, low):
    """Controlled powers on 8..14 targets (wider than one tile-pass op) run
    as matrix-core GEMMs over the groups whose control is set, with the
    squaring chain on the device: scattered targets and controls (or targets
    on the low qubits, low=True), zero exponents skipped, products of squares,
    partial and full 64-group tiles, against the oracle."""
    rng = np.random.default_rng(40 + m + n)
    qs = [int(q) for q in rng.permutation(n)[: m + len(exps)]]
    if low:
        qs = list(range(m)) + [int(q) for q in rng.permutation(np.arange(m, n))[: len(exps)]]
    d = 1 << m
    if m <= 10:
        M = haar(d, rng)
    else:   # any square matrix is a valid operand; Gaussian rows are cheap at this size
        M = (rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))) / np.sqrt(2 * d)
    p = Program()
    p.add_gates(random_circuit(n, 20, m))
    p.add_controlled_powers(qs[m:], exps, qs[:m], M)
    init = random_state(n, rng)
    want = O.program_statevector(p.kernels(), n, init=init)
    got = _sim(precision=precision).program_statevector(p, n, initial_state=init)
    tol = ATOL if precision == "complex128" else 2e-6 * max(1.0, float(np.abs(want).max()) * np.sqrt(1 << n))
    assert np.abs(got - want).max() < tol
