"""Transition columns U† H U |s> (TransitionColumnProvider,
src/qarpx/src/simulator/transition_columns.cu:236-380).

CPU: the oracle's columns (statevector of U, the Pauli sum, statevector of
U†, per basis state) equal U† H U assembled from a dense H built with
Kronecker products, and form a Hermitian matrix; the provider's validation
keeps the reference's messages.  GPU (marked): batched columns, sector
gathering and compaction against the oracle."""
import numpy as np
import pytest

import _oracle as O
from openqarp_amd import workloads as W
from openqarp_amd._abi import pack_observable, pack_stream
from openqarp_amd.errors import CapabilityError
from openqarp_amd.ir import Command, GateType, Param
from test_gpu_parity import random_circuit

PAULI = {"X": np.array([[0, 1], [1, 0]], complex), "Y": np.array([[0, -1j], [1j, 0]]),
         "Z": np.array([[1, 0], [0, -1]], complex)}


def dense_h(n, obs):
    H = np.zeros((1 << n, 1 << n), complex)
    for letters, c in obs:
        ops = {q: PAULI[p] for q, p in letters}
        M = np.eye(1)
        for q in reversed(range(n)):   # qubit 0 is the least significant bit
            M = np.kron(M, ops.get(q, np.eye(2)))
        H += complex(c) * M
    return H


def oracle_columns(U, n, obs, sources):
    U = [c for c in U if c.gate not in (GateType.GPhase, GateType.Barrier)]   # a global phase cancels
    dag = [c.dagger() for c in reversed(U)]
    po = pack_observable(obs)
    out = []
    for s in sources:
        psi = np.zeros(1 << n, complex)
        psi[int(s)] = 1.0
        psi = O.statevector(pack_stream(U), n, 0, 0, init=psi)
        phi = O.apply_observable(psi, n, po)
        out.append(O.statevector(pack_stream(dag), n, 0, 0, init=phi))
    return np.array(out)


def workload(n, seed):
    U = random_circuit(n, 30, seed) + W.brickwork_haar(n, 2, seed)
    obs = [(t, float(np.real(c))) for t, c in W.random_pauli_sum(n, 15, seed + 1)]
    return U, obs


def test_oracle_columns_are_u_dagger_h_u():
    n = 5
    U, obs = workload(n, 3)
    cols = oracle_columns(U, n, obs, range(1 << n))          # row s = U† H U e_s
    Umat = np.array([O.statevector(pack_stream(U), n, 0, 0, init=np.eye(1 << n)[s]) for s in range(1 << n)]).T
    want = Umat.conj().T @ dense_h(n, obs) @ Umat
    assert np.abs(cols.T - want).max() < 1e-12
    assert np.abs(cols - cols.conj().T).max() < 1e-12        # Hermitian


@pytest.mark.parametrize("kw,exc,msg", [
    (dict(U=[Command(GateType.Rx, 0, params=[Param.symbol("a")])]), ValueError,
     "Column commands require finite concrete parameters"),
    (dict(U=[Command(GateType.Rx, 0, params=[float("nan")])]), ValueError,
     "Column commands require finite concrete parameters"),
    (dict(U=[Command(GateType.Measure, 0, cbits=[0])]), CapabilityError,
     "Columns require unconditional measurement-free unitaries"),
    (dict(U=[Command(GateType.Reset, 0)]), CapabilityError, "Columns require unconditional measurement-free unitaries"),
    (dict(U=[Command(GateType.X, 5)]), ValueError, "TransitionColumnProvider: U acts outside the register"),
    (dict(obs=[([(0, "Z")], 1j)]), ValueError, "Column observable coefficients must be finite and real"),
    (dict(obs=[([(7, "Z")], 1.0)]), ValueError, "TransitionColumnProvider: Hamiltonian acts outside the register"),
    (dict(obs=[([(0, "Z"), (0, "X")], 1.0)]), ValueError, "TransitionColumnProvider Pauli term repeats a qubit"),
    (dict(cap=0), ValueError, "TransitionColumnProvider requires positive batch capacity"),
    (dict(sector=[3, 2]), ValueError, "TransitionColumnProvider sector targets must be strictly increasing"),
    (dict(sector=[40]), ValueError, "TransitionColumnProvider sector target is outside the register"),
])
def test_provider_validation_messages(kw, exc, msg):
    from openqarp_amd.columns import TransitionColumnProvider
    with pytest.raises(exc) as e:
        TransitionColumnProvider(None, kw.get("U", []), 4, kw.get("obs", []), kw.get("cap", 1), kw.get("sector"))
    assert str(e.value) == msg


@pytest.mark.gpu
@pytest.mark.parametrize("precision,cap", [("complex128", 3), ("complex128", 1), ("complex64", 8)])
def test_columns_against_oracle(precision, cap):
    """Seven sources in batches of the capacity (a partial last batch), every
    column against the oracle; a global phase in U cancels."""
    from openqarp_amd.simulator import QarpSimulator
    n = 8
    U, obs = workload(n, 5)
    U = [Command(GateType.GPhase, params=[0.7])] + U
    src = np.array([0, 5, 17, 255, 128, 64, 9], dtype=np.uint64)
    sim = QarpSimulator(precision=precision)
    prov = sim.make_transition_column_provider(U, n, [(t, complex(c)) for t, c in obs], cap)
    got = prov.columns(src)
    want = oracle_columns(U, n, obs, src)
    assert got.shape == (len(src), 1 << n)
    assert np.abs(got - want).max() < (1e-12 if precision == "complex128" else 2e-5)
    d = prov.diagnostics
    assert d["columns"] == len(src) and d["column_batches"] == -(-len(src) // d["capacity"])


@pytest.mark.gpu
def test_compact_columns_with_sector():
    from openqarp_amd.simulator import QarpSimulator
    n = 7
    U, obs = workload(n, 9)
    sector = np.array([1, 4, 9, 33, 70, 100, 127], dtype=np.uint64)
    sim = QarpSimulator()
    prov = sim.make_transition_column_provider(U, n, [(t, complex(c)) for t, c in obs], 4, sector)
    src = sector[[0, 2, 5]]
    offsets, labels, values = prov.compact_columns(src, 1e-3)
    want = oracle_columns(U, n, obs, src)[:, sector.astype(np.int64)]
    assert offsets[0] == 0 and len(offsets) == len(src) + 1
    for j in range(len(src)):
        keep = np.abs(want[j]) > 1e-3
        lo, hi = int(offsets[j]), int(offsets[j + 1])
        assert np.array_equal(labels[lo:hi], sector[keep])
        assert keep.any() and np.abs(values[lo:hi] - want[j][keep]).max() < 1e-12
    with pytest.raises(ValueError, match="outside the configured sector"):
        prov.compact_columns(np.array([2], dtype=np.uint64), 0.0)
    with pytest.raises(ValueError, match="finite, non-negative compaction tolerance"):
        prov.compact_columns(src, -1.0)
    empty = sim.make_transition_column_provider(U, n, [], 2)
    assert not np.any(empty.columns(np.array([3], dtype=np.uint64)))


@pytest.mark.gpu
@pytest.mark.parametrize("tol", [0.0, 0.05])
def test_compact_columns_on_device(tol):
    """Without a sector the columns are compacted on the device
    (qhip_compact): labels ascending per source, values to 1e-12, offsets
    per source, across batches."""
    from openqarp_amd.simulator import QarpSimulator
    n = 7
    U, obs = workload(n, 4)
    sim = QarpSimulator()
    prov = sim.make_transition_column_provider(U, n, [(t, complex(c)) for t, c in obs], 2)
    src = np.array([3, 100, 7, 64, 127], dtype=np.uint64)
    offsets, labels, values = prov.compact_columns(src, tol)
    want = oracle_columns(U, n, obs, src)
    assert len(offsets) == len(src) + 1 and offsets[0] == 0
    for j in range(len(src)):
        keep = np.abs(want[j]) > tol
        lo, hi = int(offsets[j]), int(offsets[j + 1])
        assert np.array_equal(labels[lo:hi], np.nonzero(keep)[0].astype(np.uint64))
        assert np.abs(values[lo:hi] - want[j][keep]).max() < 1e-12
    pruned = np.abs(want)[np.abs(want) <= tol]
    assert prov.diagnostics["max_pruned_amplitude"] == pytest.approx(pruned.max() if pruned.size else 0.0,
                                                                          abs=1e-12)
    # Host buffers follow the kept entries: at most one batch of slack.
    owner = labels
    while owner.base is not None:
        owner = owner.base
    assert owner.size <= len(labels) + 2 * (1 << n) + len(labels) // 20 + 1


@pytest.mark.gpu
def test_compaction_rejects_nonfinite_columns():
    """Two finite coefficients whose sum overflows make the column non-finite
    (inf, then NaN through U†): the device compaction raises, as the
    reference does for every entry, pruned ones included."""
    from openqarp_amd.simulator import QarpSimulator
    n = 6
    U, _ = workload(n, 2)
    obs = [([(0, "Z")], 1e308), ([(1, "Z")], 1e308)]
    prov = QarpSimulator().make_transition_column_provider(U, n, obs, 2)
    for tol in (0.0, 1e300):
        with pytest.raises(RuntimeError, match="nonfinite coefficient"):
            prov.compact_columns(np.array([0, 5], dtype=np.uint64), tol)
