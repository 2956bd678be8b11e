"""PauliAveraging's host side: the grouping strategies (the contracts the
reference's tests/test_operators/test_grouping.py pins, restated with their
known answers) and the diagonalising Clifford, checked against what fixes it
mathematically — after the emitted gates C, every term P_r of the group
satisfies C P_r C† = (-1)^{sign_r} Z^{mask_r}, as dense matrices (the check of
src/qarpx/tests/test_synthesis_pauli_exp_clifford.cpp:64-79), plus the
single-gate conjugation rules of its tableau.  No GPU calls."""
import random

import numpy as np
import pytest

import _fixtures as F
import _oracle as O
from openqarp_amd._abi import pack_stream
from openqarp_amd.grouping import (FullyCommuting, GroupingStrategy, NoGrouping, QubitWiseCommuting,
                                   canonical_term_order, commutes, compute_basis_change_clifford,
                                   diagonalise_group, greedy_first_fit, group_basis, qubit_wise_commute)
from openqarp_amd.ir import GateType
from openqarp_amd.primitives import PauliAveraging

LETTERS = "XYZ"


def _random_terms(n_qubits, n_terms, seed):
    rng = random.Random(seed)
    terms = []
    for _ in range(n_terms):
        support = rng.sample(range(n_qubits), rng.randint(1, n_qubits))
        terms.append({q: rng.choice(LETTERS) for q in support})
    return terms


def _dense(term, n):
    return F.pauli_matrix([([(q, l) for q, l in term.items()], 1.0)], n)


def _unitary(cmds, n):
    cols = []
    for j in range(1 << n):
        e = np.zeros(1 << n, np.complex128)
        e[j] = 1.0
        cols.append(O.statevector(pack_stream(cmds), n, 0, 0, init=e))
    return np.array(cols).T


@pytest.mark.parametrize("strategy", [NoGrouping(), QubitWiseCommuting(), FullyCommuting()])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_group_is_a_deterministic_partition(strategy, seed):
    terms = _random_terms(6, 40, seed)
    groups = strategy.group(terms, 6)
    assert sorted(i for g in groups for i in g) == list(range(len(terms)))
    assert groups == strategy.group(terms, 6)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_groups_commute_as_their_strategy_promises(seed):
    terms = _random_terms(6, 40, seed)
    for grp in QubitWiseCommuting().group(terms, 6):
        assert all(qubit_wise_commute(terms[i], terms[j]) for i in grp for j in grp)
        basis = group_basis(grp, terms)
        assert all(basis[q] == p for i in grp for q, p in terms[i].items())
    for grp in FullyCommuting().group(terms, 6):
        assert all(commutes(terms[i], terms[j]) for i in grp for j in grp)


def test_commutes_is_the_matrix_commutator():
    terms = _random_terms(4, 30, 11)
    for a in terms[:12]:
        for b in terms[:12]:
            A, B = _dense(a, 4), _dense(b, 4)
            assert commutes(a, b) == np.allclose(A @ B, B @ A)


def test_known_answers_of_the_reference_grouping_tests():
    terms = [{0: "X", 1: "X"}, {0: "Y", 1: "Y"}]
    assert FullyCommuting().group(terms, 2) == [[0, 1]]
    assert QubitWiseCommuting().group(terms, 2) == [[0], [1]]
    assert not FullyCommuting.qubit_wise and QubitWiseCommuting.qubit_wise and NoGrouping.qubit_wise
    terms = [{1: "Y", 2: "Y"}, {0: "Z"}, {0: "X", 1: "X"}, {1: "X", 2: "X"}, {0: "Y", 1: "Y"}]
    assert canonical_term_order(terms) == [1, 2, 4, 3, 0]
    compat = {(1, 0): True, (2, 0): True, (2, 1): False}
    assert greedy_first_fit(3, lambda i, j: compat.get((i, j), compat.get((j, i), True))) == [[0, 1], [2]]
    assert NoGrouping().group(_random_terms(4, 6, 3), 4) == [[i] for i in range(6)]
    for strategy in (NoGrouping(), QubitWiseCommuting(), FullyCommuting()):
        assert strategy.group([], 3) == []
    assert repr(FullyCommuting()) == "FullyCommuting()" and isinstance(NoGrouping(), GroupingStrategy)


@pytest.mark.parametrize("strategy", [QubitWiseCommuting(), FullyCommuting()])
@pytest.mark.parametrize("seed", [0, 1, 2, 3, 4])
def test_grouping_is_order_insensitive(strategy, seed):
    terms = _random_terms(5, 20, 99)

    def as_sets(groups, ts):
        return sorted(sorted(tuple(sorted(ts[i].items())) for i in g) for g in groups)

    shuffled = terms[:]
    random.Random(seed).shuffle(shuffled)
    assert as_sets(strategy.group(shuffled, 5), shuffled) == as_sets(strategy.group(terms, 5), terms)


def test_tableau_rules_of_single_gates():
    """H: X <-> Z, Y -> -Y; Sdg: Y -> X, X -> -Y; the emitted gates bring each
    single-qubit Pauli to +-Z with the sign the conjugation gives."""
    cl, z, s = compute_basis_change_clifford(["X"], 1)
    assert [c.gate for c in cl] == [GateType.H] and z == [[True]] and s == [False]
    cl, z, s = compute_basis_change_clifford(["Y"], 1)
    assert [c.gate for c in cl] == [GateType.Sdg, GateType.H] and z == [[True]] and s == [False]
    cl, z, s = compute_basis_change_clifford(["Z"], 1)
    assert cl == [] and z == [[True]] and s == [False]
    # XX -> Z0 (H, H, CX(1, 0)); ZZ becomes X1, then H -> Z1; YY = -XX.ZZ -> -Z0Z1.
    cl, z, s = compute_basis_change_clifford(["XX", "YY", "ZZ"], 2)
    assert [(c.gate, c.qubits) for c in cl] == [(GateType.H, (0,)), (GateType.H, (1,)), (GateType.CX, (1, 0)),
                                                (GateType.H, (1,))]
    assert z == [[True, False], [True, True], [False, True]] and s == [False, True, False]
    cl, z, s = compute_basis_change_clifford([], 2)
    assert cl == [] and z == [] and s == []
    with pytest.raises(ValueError, match="does not commute"):
        compute_basis_change_clifford(["X", "Z"], 1)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_clifford_diagonalises_every_fully_commuting_group(n, seed):
    """C P C† = s Z^mask for every term of every group, as dense matrices."""
    terms = _random_terms(n, 3 * n + 4, 100 * n + seed)
    for grp in FullyCommuting().group(terms, n):
        cl, masks, signs = diagonalise_group([terms[i] for i in grp], n)
        U = _unitary(cl, n)
        assert np.allclose(U.conj().T @ U, np.eye(1 << n), atol=1e-12)
        for i, m, sg in zip(grp, masks, signs):
            want = np.diag([(-1.0 if sg else 1.0) * (-1.0) ** bin(k & m).count("1") for k in range(1 << n)])
            assert np.abs(U @ _dense(terms[i], n) @ U.conj().T - want).max() < 1e-12


def test_pauli_averaging_groups_and_known_answers():
    """Counts of groups (XX, YY, ZZ commute: one group; the reference's
    test_pauli_averaging.py:59-90), the constant, and the parity estimator on
    the all-zero and all-one outcomes."""
    from openqarp_amd.ir import SamplingResult
    obs = [([(0, "X"), (1, "X")], 1.0), ([(0, "Y"), (1, "Y")], 1.0), ([(0, "Z"), (1, "Z")], 1.0)]
    pa = PauliAveraging([], 2, obs, 100)
    assert pa.n_terms == 3 and pa.n_groups == 1 and isinstance(pa.grouping, FullyCommuting)
    assert PauliAveraging([], 2, obs, 100, QubitWiseCommuting()).n_groups == 3
    assert PauliAveraging([], 2, obs, 100, NoGrouping()).n_groups == 3
    pa = PauliAveraging([], 2, [([(0, "Z")], 1.0), ([], 0.5)], 10)
    assert pa.constant == 0.5 and pa.n_groups == 1
    zero = SamplingResult(n_qubits=2, n_shots=10, counts={0: 10})
    assert pa.group_estimate(zero, 0) == 1.0
    one = SamplingResult(n_qubits=2, n_shots=10, counts={1: 10})
    assert pa.group_estimate(one, 0) == -1.0
    assert PauliAveraging([], 2, [([], 3.0)], 10).n_groups == 0
    assert "FullyCommuting" in repr(pa)
