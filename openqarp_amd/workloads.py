"""Seeded circuit and observable builders for the benchmark workloads.

Each builder returns plain ``Command`` streams and observables in the
simulator ABI, constructed the way the reference's blocks emit them:
  * hardware-efficient ansatz: RY on every qubit, RZ on every qubit, CX
    ladder q -> q+1 (src/qarp/blocks/_primitives/hea_block.py:64-85);
  * brickwork CZ entangler: even bonds then odd bonds
    (blocks/_primitives/brickwork_entangling_block.py:42-56);
  * Haar-random unitaries: complex Ginibre, QR, phase-fixed diagonal
    (blocks/_primitives/haar_random_block.py:169-176);
  * QAOA: RZZ(w*gamma) per edge then RX(beta) per qubit
    (blocks/_primitives/cost_operator_block.py:47-62,
    mixed_operator_block.py:33-36).
All randomness comes from ``numpy.random.default_rng(seed)``.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Optional

import numpy as np

from .ir import Command, GateType, Param

TWO_PI = 2.0 * math.pi


def haar_unitary(dim: int, rng: np.random.Generator) -> np.ndarray:
    z = (rng.standard_normal((dim, dim)) + 1j * rng.standard_normal((dim, dim))) / np.sqrt(2.0)
    q, r = np.linalg.qr(z)
    d = np.diagonal(r)
    return q * (d / np.abs(d))[np.newaxis, :]


@dataclass
class Workload:
    n_qubits: int
    commands: list
    observable: list
    values: Optional[dict] = None       # symbol -> value for symbolic streams
    symbols: Optional[list] = None      # gradient order


def hea(n: int, layers: int, seed: int, symbolic: bool = False) -> Workload:
    """RY(theta) then RZ(phi) on every qubit, then CX(q, q+1), per layer."""
    rng = np.random.default_rng(seed)
    cmds, values, order = [], {}, []
    for layer in range(layers):
        th = rng.uniform(0.0, TWO_PI, n)
        ph = rng.uniform(0.0, TWO_PI, n)
        for gate, angles, tag in ((GateType.Ry, th, "theta"), (GateType.Rz, ph, "phi")):
            for q in range(n):
                name = f"{tag}_{layer}_{q}"
                if symbolic:
                    values[name] = float(angles[q])
                    order.append(name)
                    cmds.append(Command(gate, q, Param.symbol(name)))
                else:
                    cmds.append(Command(gate, q, float(angles[q])))
        for q in range(n - 1):
            cmds.append(Command(GateType.CX, q, q + 1))
    return Workload(n, cmds, [], values or None, order or None)


def tfim_chain(n: int, field: float = 0.7, zz: float = 1.0) -> list:
    """sum_i zz Z_i Z_{i+1} + field sum_i X_i on an open chain."""
    obs = [([(i, "Z"), (i + 1, "Z")], complex(zz)) for i in range(n - 1)]
    obs += [([(i, "X")], complex(field)) for i in range(n)]
    return obs


def c1_hea(n: int = 20, layers: int = 10, seed: int = 1) -> Workload:
    w = hea(n, layers, seed)
    w.observable = tfim_chain(n, 0.7)
    return w


def brickwork_haar(n: int, layers: int, seed: int) -> list:
    """Per layer a seeded random perfect matching of the qubits, each pair
    carrying a Haar-random 2-qubit unitary (as a Custom command)."""
    rng = np.random.default_rng(seed)
    cmds = []
    for _ in range(layers):
        perm = rng.permutation(n)
        for i in range(0, n - 1, 2):
            a, b = int(perm[i]), int(perm[i + 1])
            cmds.append(Command(GateType.Custom, qubits=[a, b], unitary=haar_unitary(4, rng)))
    return cmds


def random_pauli_sum(n: int, n_terms: int, seed: int,
                     probs=(0.7, 0.1, 0.1, 0.1)) -> list:
    """Each qubit drawn I/X/Y/Z with the given probabilities; N(0,1) coefficients."""
    rng = np.random.default_rng(seed)
    letters = np.array(["I", "X", "Y", "Z"])
    obs = []
    for _ in range(n_terms):
        draw = rng.choice(4, size=n, p=probs)
        pauli = [(q, str(letters[d])) for q, d in enumerate(draw) if d != 0]
        obs.append((pauli, complex(rng.standard_normal())))
    return obs


def c2_brickwork(n: int = 30, layers: int = 30, seed: int = 2, n_terms: int = 2000) -> Workload:
    return Workload(n, brickwork_haar(n, layers, seed), random_pauli_sum(n, n_terms, seed + 1))


def brickwork_cz_pairs(n: int) -> list:
    even = [(2 * i, 2 * i + 1) for i in range(n // 2)]
    odd = [(2 * i + 1, 2 * i + 2) for i in range((n - 1) // 2)]
    return even + odd


def c3_ansatz(n: int = 24, layers: int = 8, seed: int = 3) -> Workload:
    """RY and RZ on every qubit (symbolic), then CZ brickwork, per layer."""
    rng = np.random.default_rng(seed)
    cmds, values, order = [], {}, []
    for layer in range(layers):
        for gate, tag in ((GateType.Ry, "ry"), (GateType.Rz, "rz")):
            for q in range(n):
                name = f"{tag}_{layer}_{q}"
                values[name] = float(rng.uniform(0.0, TWO_PI))
                order.append(name)
                cmds.append(Command(gate, q, Param.symbol(name)))
        for a, b in brickwork_cz_pairs(n):
            cmds.append(Command(GateType.CZ, a, b))
    return Workload(n, cmds, [], values, order)


def random_local_pauli_sum(n: int, n_terms: int, seed: int, max_weight: int = 4) -> list:
    """Weight uniform in 1..max_weight on distinct uniformly chosen qubits,
    letters uniform in X/Y/Z, coefficients N(0,1)/sqrt(n_terms)."""
    rng = np.random.default_rng(seed)
    obs = []
    for _ in range(n_terms):
        w = int(rng.integers(1, max_weight + 1))
        qs = sorted(int(q) for q in rng.choice(n, size=w, replace=False))
        ls = rng.choice(["X", "Y", "Z"], size=w)
        obs.append(([(q, str(l)) for q, l in zip(qs, ls)],
                    complex(rng.standard_normal() / math.sqrt(n_terms))))
    return obs


def c3_gradient(n: int = 24, layers: int = 8, seed: int = 3, n_terms: int = 1000) -> Workload:
    w = c3_ansatz(n, layers, seed)
    w.observable = random_local_pauli_sum(n, n_terms, seed + 1)
    return w


def random_regular_graph(d: int, n: int, seed: int) -> list:
    """Seeded random d-regular simple graph by the pairing model with
    rejection (uniform over d-regular multigraph pairings, restarted on a
    loop or repeated edge)."""
    rng = np.random.default_rng(seed)
    while True:
        stubs = np.repeat(np.arange(n), d)
        rng.shuffle(stubs)
        pairs = stubs.reshape(-1, 2)
        edges = set()
        ok = True
        for a, b in pairs:
            a, b = int(min(a, b)), int(max(a, b))
            if a == b or (a, b) in edges:
                ok = False
                break
            edges.add((a, b))
        if ok:
            return sorted(edges)


def qaoa_maxcut(n: int, p: int, seed: int, symbolic: bool = False) -> Workload:
    """H on all qubits, then p layers of RZZ(w*gamma_l) per edge and
    RX(beta_l) per qubit (cost_operator_block.py:47-62, mixed_operator_block.py:
    33-36); an odd n uses the (n+1)-node graph minus its last node; <C> = sum over edges of Z_a Z_b (the Ising cost
    form of graph_to_cost_hamiltonian, src/qarp/graphs/_utils.py:13-50)."""
    rng = np.random.default_rng(seed)
    if (3 * n) % 2:
        # No 3-regular graph has an odd number of nodes: take the graph on
        # n + 1 nodes from the same seed and drop node n with its edges.
        edges = [e for e in random_regular_graph(3, n + 1, seed) if n not in e]
    else:
        edges = random_regular_graph(3, n, seed)
    gammas = rng.uniform(0.0, math.pi, p)
    betas = rng.uniform(0.0, math.pi / 2, p)
    cmds = [Command(GateType.H, q) for q in range(n)]
    values, order = {}, []
    for layer in range(p):
        g = Param.symbol(f"gamma_{layer}") if symbolic else float(gammas[layer])
        b = Param.symbol(f"beta_{layer}") if symbolic else float(betas[layer])
        values[f"gamma_{layer}"] = float(gammas[layer])
        values[f"beta_{layer}"] = float(betas[layer])
        order += [f"gamma_{layer}", f"beta_{layer}"]
        for a, c in edges:
            cmds.append(Command(GateType.RZZ, a, c, g * 1.0 if symbolic else g))
        for q in range(n):
            cmds.append(Command(GateType.Rx, q, b))
    obs = [([(a, "Z"), (c, "Z")], 1.0 + 0j) for a, c in edges]
    return Workload(n, cmds, obs, values if symbolic else None, order if symbolic else None)


def batched_vqe(n_instances: int = 512, seed: int = 4, n_min: int = 12, n_max: int = 22,
                d_min: int = 4, d_max: int = 16) -> list:
    """Independent HEA instances with their own TFIM (h uniform in [0.2, 2.0])."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_instances):
        n = int(rng.integers(n_min, n_max + 1))
        depth = int(rng.integers(d_min, d_max + 1))
        h = float(rng.uniform(0.2, 2.0))
        w = hea(n, depth, int(rng.integers(0, 2**31 - 1)), symbolic=True)
        w.observable = tfim_chain(n, h, 1.0)
        out.append(w)
    return out


def structured_program(n: int = 12, seed: int = 8):
    """A seeded structured program exercising every kernel kind: brickwork
    gates, two basis permutations (the second folds into the first), a dense
    3-qubit block, QPE-style controlled powers of a 2-qubit Haar unitary
    (exponents 2^j on the top qubits, plus a zero and a multi-bit exponent),
    then more gates."""
    from .program import Program
    rng = np.random.default_rng(seed)
    p = Program()
    p.add_gates(brickwork_haar(n, 3, seed))
    q1 = [int(q) for q in rng.choice(n, size=min(n, 5), replace=False)]
    p.add_permutation(q1, rng.permutation(1 << len(q1)))
    q2 = [int(q) for q in rng.choice(n, size=min(n, 3), replace=False)]
    p.add_permutation(q2, rng.permutation(1 << len(q2)))
    p.add_dense([int(q) for q in rng.choice(n, size=3, replace=False)], haar_unitary(8, rng))
    controls = list(range(n - 1, max(n - 6, 2) - 1, -1))
    exps = [1 << j for j in range(len(controls))]
    exps[1] = 0
    exps[-1] = 11
    p.add_controlled_powers(controls, exps, [0, 1], haar_unitary(4, rng))
    p.add_gates(brickwork_haar(n, 2, seed + 1))
    return p
