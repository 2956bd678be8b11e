"""Partition of Pauli terms into commuting groups, and the Clifford that
diagonalises a group — the host side of PauliAveraging.

Follows the reference's grouping contract (src/qarp/operators/_grouping.py):
terms are sparse ``{qubit: 'X'|'Y'|'Z'}`` dicts; a strategy returns index
groups only, never coefficients; QubitWiseCommuting and FullyCommuting run a
greedy first fit over ``canonical_term_order`` (support, then letters), so a
permuted Hamiltonian gives the permuted partition (:99-142, 195-226);
NoGrouping keeps one term per group in input order (:178-184).

``compute_basis_change_clifford`` restates the reference's symplectic
elimination (src/qarpx/src/synthesis/pauli_exponential.cpp:13-186): the group
as an x/z/sign tableau; for each row in order, the first unprocessed qubit
where it is not I becomes its pivot; the row is turned to Z there (Sdg where
it is Y, then H where it is X), every other unprocessed qubit where it is not
I is turned to Z the same way and folded into the pivot by CX(qubit, pivot);
the pivot is then processed.  Every emitted gate updates the tableau with its
conjugation rule and sign (H: Y -> -Y; Sdg: X -> -Y; CX: the
Aaronson-Gottesman phase x_c z_t (x_t xor z_c xor 1)).  Afterwards row r is
``(-1)^signs[r] Z^{z_rows[r]}``.
"""
from __future__ import annotations

from typing import Callable, Dict, List, Sequence, Tuple

import numpy as np

from .ir import Command, GateType

PauliDict = Dict[int, str]


# ── helpers ─────────────────────────────────────────────────────────────────

def qubit_wise_commute(p1: PauliDict, p2: PauliDict) -> bool:
    """Every shared qubit carries the same letter."""
    return all(p1[q] == p2[q] for q in p1.keys() & p2.keys())


def commutes(p1: PauliDict, p2: PauliDict) -> bool:
    """Two Pauli strings commute iff they differ on an even number of the
    qubits where both are not the identity."""
    return sum(p1[q] != p2[q] for q in p1.keys() & p2.keys()) % 2 == 0


def group_basis(group_indices: Sequence[int], terms: Sequence[PauliDict]) -> Dict[int, str]:
    """Per-qubit letter of a qubit-wise commuting group (the union of its terms)."""
    basis: Dict[int, str] = {}
    for i in group_indices:
        basis.update(terms[i])
    return basis


def term_mask(pauli: PauliDict) -> int:
    m = 0
    for q in pauli:
        m |= 1 << int(q)
    return m


def canonical_term_order(terms: Sequence[PauliDict]) -> List[int]:
    """Indices of ``terms`` sorted by (sorted support, letters on it); a stable
    sort, so exact duplicates keep their input order."""
    def key(i):
        qs = sorted(terms[i])
        return tuple(qs), tuple(terms[i][q] for q in qs)
    return sorted(range(len(terms)), key=key)


def greedy_first_fit(n_items: int, compatible: Callable[[int, int], bool]) -> List[List[int]]:
    """Item i joins the first group all of whose members are compatible with
    it, else opens a new group."""
    groups: List[List[int]] = []
    for i in range(n_items):
        for grp in groups:
            if all(compatible(i, j) for j in grp):
                grp.append(i)
                break
        else:
            groups.append([i])
    return groups


def _grouped_in_canonical_order(terms: Sequence[PauliDict], compatible) -> List[List[int]]:
    perm = canonical_term_order(terms)
    groups = greedy_first_fit(len(terms), lambda i, j: compatible(perm[i], perm[j]))
    return [[perm[i] for i in grp] for grp in groups]


# ── strategies ──────────────────────────────────────────────────────────────

class GroupingStrategy:
    """Partitions Pauli terms into groups measured by one circuit each."""

    qubit_wise = False

    def group(self, terms: Sequence[PauliDict], n_qubits: int) -> List[List[int]]:
        raise NotImplementedError

    def __repr__(self) -> str:
        return f"{type(self).__name__}()"


class NoGrouping(GroupingStrategy):
    """One term per group, in input order."""

    qubit_wise = True

    def group(self, terms, n_qubits):
        return [[i] for i in range(len(terms))]


class QubitWiseCommuting(GroupingStrategy):
    """Greedy first fit under qubit-wise commutation (per-qubit basis change)."""

    qubit_wise = True

    def group(self, terms, n_qubits):
        return _grouped_in_canonical_order(terms, lambda i, j: qubit_wise_commute(terms[i], terms[j]))


class FullyCommuting(GroupingStrategy):
    """Greedy first fit under general commutation (an entangling Clifford
    diagonalises each group); the reference's default."""

    qubit_wise = False

    def group(self, terms, n_qubits):
        return _grouped_in_canonical_order(terms, lambda i, j: commutes(terms[i], terms[j]))


# ── diagonalising Clifford ──────────────────────────────────────────────────

def _to_dict(p) -> PauliDict:
    if isinstance(p, dict):
        return {int(q): str(l) for q, l in p.items() if str(l) != "I"}
    if isinstance(p, str):
        return {q: l for q, l in enumerate(p) if l != "I"}
    return {int(q): str(l) for q, l in p if str(l) != "I"}


def _label(p: PauliDict, n: int) -> str:
    return "".join(p.get(q, "I") for q in range(n))


def compute_basis_change_clifford(paulis: Sequence, n_qubits: int) -> Tuple[List[Command], List[List[bool]],
                                                                             List[bool]]:
    """(clifford commands, z_rows, signs) for mutually commuting Paulis
    (dicts, ``'IXYZ'`` strings indexed by qubit, or (qubit, letter) lists):
    after the commands, Pauli r equals (-1 if signs[r] else 1) times Z on the
    qubits where z_rows[r] is True.  ValueError if two of them do not commute."""
    rows = [_to_dict(p) for p in paulis]
    n = int(n_qubits)
    for r, p in enumerate(rows):
        if any(q < 0 or q >= n for q in p) or any(l not in "XYZ" for l in p.values()):
            raise ValueError(f"compute_basis_change_clifford: paulis[{r}] does not fit {n} qubits")
    for i in range(len(rows)):
        for j in range(i + 1, len(rows)):
            if not commutes(rows[i], rows[j]):
                raise ValueError(f"compute_basis_change_clifford: paulis[{i}] = {_label(rows[i], n)} does not "
                                 f"commute with paulis[{j}] = {_label(rows[j], n)}")
    m = len(rows)
    x = np.zeros((m, n), dtype=bool)
    z = np.zeros((m, n), dtype=bool)
    sign = np.zeros(m, dtype=bool)
    for r, p in enumerate(rows):
        for q, l in p.items():
            x[r, q] = l in "XY"
            z[r, q] = l in "YZ"
    out: List[Command] = []

    def h(q):
        sign[:] ^= x[:, q] & z[:, q]                 # H Y H = -Y
        t = x[:, q].copy()
        x[:, q] = z[:, q]
        z[:, q] = t
        out.append(Command(GateType.H, q))

    def sdg(q):
        sign[:] ^= x[:, q] & ~z[:, q]                # Sdg X S = -Y
        z[:, q] ^= x[:, q]
        out.append(Command(GateType.Sdg, q))

    def cx(c, t):
        xc, zc, xt, zt = x[:, c].copy(), z[:, c].copy(), x[:, t].copy(), z[:, t].copy()
        sign[:] ^= xc & zt & ~(xt ^ zc)
        x[:, t] = xt ^ xc
        z[:, c] = zc ^ zt
        out.append(Command(GateType.CX, c, t))

    def to_z(r, q):
        if x[r, q] and z[r, q]:
            sdg(q)
        if x[r, q]:
            h(q)

    done = np.zeros(n, dtype=bool)
    for r in range(m):
        cand = np.nonzero(~done & (x[r] | z[r]))[0]
        if cand.size == 0:
            continue
        pivot = int(cand[0])
        to_z(r, pivot)
        for q in range(n):
            if q == pivot or done[q] or not (x[r, q] or z[r, q]):
                continue
            to_z(r, q)
            cx(q, pivot)
        done[pivot] = True
    return out, [[bool(b) for b in z[r]] for r in range(m)], [bool(s) for s in sign]


def diagonalise_group(group_paulis: Sequence, n_qubits: int) -> Tuple[List[Command], List[int], List[bool]]:
    """(clifford commands, z masks as integers, signs) of a commuting group."""
    clifford, z_rows, signs = compute_basis_change_clifford(group_paulis, n_qubits)
    masks = [sum(1 << q for q, bit in enumerate(row) if bit) for row in z_rows]
    return clifford, masks, signs
