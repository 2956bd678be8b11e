"""Loading of the golden fixtures (tests/golden/*.npz, written by
tests/golden/make_golden.py from the reference simulator)."""
from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden"


def load(name):
    return np.load(GOLDEN / name, allow_pickle=False)


def observable(f):
    qs, ls = f["obs_qubits"], f["obs_letters"]
    obs = []
    for off, ln, c in zip(f["obs_offsets"], f["obs_lens"], f["obs_coeffs"]):
        pauli = [(int(qs[off + i]), chr(int(ls[off + i]))) for i in range(int(ln))]
        obs.append((pauli, complex(c)))
    return obs


def pauli_matrix(obs, n):
    """Dense 2^n x 2^n matrix of a Pauli sum by explicit Kronecker products
    (LSB: qubit 0 is the innermost factor)."""
    P = {"I": np.eye(2), "X": np.array([[0, 1], [1, 0]]), "Y": np.array([[0, -1j], [1j, 0]]),
         "Z": np.diag([1.0, -1.0])}
    H = np.zeros((1 << n, 1 << n), dtype=np.complex128)
    for pauli, c in obs:
        letters = ["I"] * n
        for q, l in pauli:
            letters[q] = l
        m = np.array([[1.0 + 0j]])
        for q in reversed(range(n)):
            m = np.kron(m, P[letters[q]])
        H += c * m
    return H
