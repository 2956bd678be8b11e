"""Planner derivations (permutation tables, local unitaries, command digests)
against the reference build, the reference's own known answers
(src/qarpx/tests/test_program_kernels.cpp:214-460) and brute force through
the dense unitary of the oracle."""
import math

import numpy as np
import pytest

import _oracle as O
from openqarp_amd import derive as D
from openqarp_amd._abi import pack_stream
from openqarp_amd.ir import Command, GateType

needs_ref = pytest.mark.skipif(O.reference() is None, reason="reference build (oracle/_ref) not present")


def dense_unitary(cmds, n):
    """Columns U|c> by the oracle statevector from each basis state."""
    ps = pack_stream(cmds)
    d = 1 << n
    U = np.zeros((d, d), dtype=np.complex128)
    for c in range(d):
        init = np.zeros(d, dtype=np.complex128)
        init[c] = 1.0
        U[:, c] = O.statevector(ps, n, 1, 0, init=init)
    return U


def permutation_of(U, tol=1e-9):
    """The table of U when it is a permutation with trivial phases, else None."""
    d = U.shape[0]
    t = np.zeros(d, dtype=np.int64)
    for c in range(d):
        y = int(np.argmax(np.abs(U[:, c])))
        col = U[:, c].copy()
        if abs(col[y] - 1.0) > tol:
            return None
        col[y] = 0
        if np.abs(col).max() > tol:
            return None
        t[c] = y
    if len(set(t.tolist())) != d:
        return None
    return t


def qft_on(reg):
    out = []
    k = len(reg)
    for j in range(k - 1, -1, -1):
        out.append(Command(GateType.H, reg[j]))
        for m in range(j - 1, -1, -1):
            out.append(Command(GateType.CP, reg[m], reg[j], math.pi / (1 << (j - m))))
    return out


def controlled_add(reg, ctrl, a):
    k = len(reg)
    q = qft_on(reg)
    out = list(q)
    for j in range(k):
        angle = 2.0 * math.pi * a * (1 << (k - 1 - j)) / (1 << k)
        out.append(Command(GateType.CP, ctrl, reg[j], angle) if ctrl is not None
                   else Command(GateType.P, reg[j], angle))
    out += [c.dagger() for c in reversed(q)]
    return out


def mcz(*qs):
    return Command(GateType.MCZ, qubits=list(qs))


def test_classical_stream_table_is_its_permutation():
    cmds = [Command(GateType.X, 1), Command(GateType.CX, 1, 3), Command(GateType.CCX, qubits=[0, 3, 2]),
            Command(GateType.SWAP, 0, 2), Command(GateType.CSWAP, qubits=[3, 1, 0]), Command(GateType.H, 2),
            mcz(0, 1, 3, 2), Command(GateType.H, 2)]
    t = D.permutation_table(cmds, [0, 1, 2, 3], 0, 0)
    assert t is not None and t.dtype == np.int64
    assert np.array_equal(t, permutation_of(dense_unitary(cmds, 4)))


def test_draper_adder_needs_its_register_simulated():
    """x -> x + 3 mod 8: QFT, phases, inverse QFT — not classical gate by gate."""
    k = 3
    cmds = controlled_add([0, 1, 2], None, 3)
    want = permutation_of(dense_unitary(cmds, k))
    assert want is not None and np.array_equal(want, (np.arange(8) + 3) % 8)
    assert D.permutation_table(cmds, [0, 1, 2], 0, 0) is None
    assert np.array_equal(D.permutation_table(cmds, [0, 1, 2], 3, 64), want)


def test_relative_phase_is_never_a_permutation():
    assert D.permutation_table([Command(GateType.X, 0), Command(GateType.Z, 0)], [0], 8, 64) is None
    assert D.permutation_table([Command(GateType.CZ, 0, 1)], [0, 1], 8, 64) is None


def test_controlled_adders_derive_with_only_the_register_simulated():
    cmds = []
    for ctrl, a in ((3, 1), (4, 3), (5, 6)):
        cmds += controlled_add([0, 1, 2], ctrl, a)
    want = permutation_of(dense_unitary(cmds, 6))
    v = np.arange(64)
    shift = ((v >> 3) & 1) * 1 + ((v >> 4) & 1) * 3 + ((v >> 5) & 1) * 6
    assert np.array_equal(want, (v & ~7) | (((v & 7) + shift) % 8))
    # 2^3 assignments x 4^3 columns: within a budget far below 4^6.
    assert np.array_equal(D.permutation_table(cmds, list(range(6)), 3, 9), want)


def test_mcx_pattern_needs_its_closing_hadamard_on_the_target():
    cmds = [Command(GateType.H, 0), mcz(1, 0), Command(GateType.H, 1)]
    assert permutation_of(dense_unitary(cmds, 2)) is None
    assert D.permutation_table(cmds, [0, 1], 2, 64) is None


def test_subset_remap_and_rejections():
    """Local bit b <-> qubits[b]; parametric, conditioned, measured or
    out-of-subset streams have no table; repeated qubits raise."""
    from openqarp_amd.ir import Param
    cmds = [Command(GateType.CX, 7, 2)]
    t = D.permutation_table(cmds, [2, 7], 0, 0)   # local 0 = qubit 2 (target), 1 = qubit 7 (control)
    assert np.array_equal(t, [0, 1, 3, 2])
    assert D.permutation_table(cmds, [2], 8, 64) is None
    assert D.permutation_table([Command(GateType.Rx, 0, Param.symbol("a"))], [0], 8, 64) is None
    assert D.permutation_table([Command(GateType.X, 0, condition_bits=[0], condition_values=[True])],
                               [0], 8, 64) is None
    assert D.permutation_table([Command(GateType.Measure, 0, cbits=[0])], [0], 8, 64) is None
    assert D.permutation_table([], [], 8, 64) is None
    with pytest.raises(ValueError, match="permutation_table: qubits must be distinct"):
        D.permutation_table(cmds, [2, 2], 0, 0)
    with pytest.raises(ValueError, match="local_unitary_of: 1..12 qubits required"):
        D.local_unitary(cmds, [])
    with pytest.raises(ValueError, match="local_unitary_of: commands must be concrete"):
        D.local_unitary(cmds, [2])


def _random_trial(rng):
    cmds = []
    for _ in range(2 + int(rng.integers(3))):
        kind = int(rng.integers(7))
        if kind == 0:
            cmds += controlled_add([0, 1, 2], 3 + int(rng.integers(2)), 1 + int(rng.integers(7)))
        elif kind == 1:
            cmds.append(Command(GateType.CCX, qubits=[3, 4, int(rng.integers(3))]))
        elif kind == 2:
            cmds.append(Command(GateType.SWAP, 3, 4))
        elif kind == 3:
            t = int(rng.integers(3))
            cmds += [Command(GateType.H, t), mcz(3, 4, t), Command(GateType.H, t),
                     Command(GateType.CP, 3, 4, math.pi), Command(GateType.CP, 3, 4, -math.pi)]
        elif kind == 4:
            cmds += controlled_add([0, 1, 2], 3, 1 + int(rng.integers(7)))
            cmds.append(Command(GateType.CX, int(rng.integers(3)), 4))
        elif kind == 5:
            theta = 0.4 + 0.1 * int(rng.integers(5))
            cmds.append(Command(GateType.GPhase, params=[theta]))
            cmds += controlled_add([0, 1, 2], None, 1 + int(rng.integers(7)))
            if rng.integers(2):
                cmds.append(Command(GateType.GPhase, params=[-theta]))
        else:
            theta = math.pi if rng.integers(2) else 0.5
            cmds += [Command(GateType.CP, 3, 4, theta), Command(GateType.CP, 3, 4, -theta)]
            if rng.integers(3) == 0:
                cmds.append(Command(GateType.Rz, int(rng.integers(5)), 0.3))
    return cmds


@needs_ref
def test_random_circuits_match_the_reference_build_and_brute_force():
    """Adders, classical gates and phases that may or may not cancel: the
    table equals the reference build's (None included) and the dense
    unitary's permutation."""
    rng = np.random.default_rng(20260928)
    pos = neg = 0
    for trial in range(120):
        cmds = _random_trial(rng)
        want = permutation_of(dense_unitary(cmds, 5))
        for max_rest, work in ((5, 64), (3, 9), (0, 0)):
            got = D.permutation_table(cmds, list(range(5)), max_rest, work)
            ref = O.ref_permutation_table(pack_stream(cmds), list(range(5)), max_rest, work)
            assert (got is None) == (ref is None), (trial, max_rest)
            if got is not None:
                assert np.array_equal(got, ref) and np.array_equal(got, want), (trial, max_rest)
        got = D.permutation_table(cmds, list(range(5)), 5, 64)
        if want is None:
            neg += 1
            assert got is None
        else:
            pos += 1
            assert got is not None
    assert pos > 20 and neg > 20


@needs_ref
def test_digests_equal_the_reference_build():
    rng = np.random.default_rng(5)
    for trial in range(20):
        cmds = [Command(GateType.H, 0), Command(GateType.Rx, 1, float(rng.uniform(0, 6))),
                Command(GateType.CX, 0, 2), Command(GateType.Measure, 2, cbits=[1]),
                Command(GateType.U, 1, params=[float(x) for x in rng.uniform(0, 6, 3)]),
                Command(GateType.Custom, qubits=[2, 0], unitary=np.linalg.qr(
                    rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4)))[0]),
                Command(GateType.GPhase, params=[float(rng.uniform(0, 6))])][: 2 + trial % 6]
        ps = pack_stream(cmds)
        assert D.commands_digest(cmds) == O.ref_commands_digest(ps)
        for qs in ([0, 1, 2], [2, 0, 1, 5], [1, 0]):
            assert D.local_commands_digest(cmds, qs) == O.ref_local_commands_digest(ps, qs), (trial, qs)
    parts = [([Command(GateType.H, 0)], 3), ([Command(GateType.CX, 0, 1), Command(GateType.X, 1)], 2)]
    flat = [Command(GateType.H, 0)] * 3 + [Command(GateType.CX, 0, 1), Command(GateType.X, 1)] * 2
    assert D.parts_digest(parts) == D.commands_digest(flat) == O.ref_commands_digest(pack_stream(flat))


@pytest.mark.gpu
def test_local_unitary_against_the_reference_build():
    """The device's 2k-qubit identity register against local_unitary_of of
    the reference build, on subsets in arbitrary order, k = 1..8 and 12."""
    if O.reference() is None:
        pytest.skip("reference build (oracle/_ref) not present")
    rng = np.random.default_rng(3)
    names = [GateType.H, GateType.Rx, GateType.Ry, GateType.CX, GateType.CZ, GateType.SWAP, GateType.T,
             GateType.CP, GateType.RZZ]
    for k in (1, 2, 3, 5, 8, 12):
        qubits = [int(q) for q in rng.permutation(16)[:k]]
        cmds = [Command(GateType.GPhase, params=[0.3])]
        for _ in range(4 * k):
            g = names[int(rng.integers(len(names)))]
            two = g in (GateType.CX, GateType.CZ, GateType.SWAP, GateType.CP, GateType.RZZ)
            if two and k < 2:
                continue
            qs = [int(q) for q in rng.choice(qubits, 2 if two else 1, replace=False)]
            params = [float(rng.uniform(0, 6))] if g in (GateType.Rx, GateType.Ry, GateType.CP, GateType.RZZ) else []
            cmds.append(Command(g, qubits=qs, params=params))
        U = D.local_unitary(cmds, qubits)
        ref = O.ref_local_unitary_of(pack_stream(cmds), qubits)
        assert U.shape == (1 << k, 1 << k)
        assert np.abs(U - ref).max() < 1e-12, k
