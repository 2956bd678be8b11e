"""qhip_transition_noalloc, the Pauli-sum expectation entry point that
allocates no device memory (its group table and partials live in a fixed
module-scope array), against qhip_transition through the C ABI: bit-equal
results on consecutive calls, the second on a smaller input, and a clear
error for observables beyond the fixed array's capacity."""
import ctypes as C

import numpy as np
import pytest

from openqarp_amd import _lib
from openqarp_amd import workloads as W
from openqarp_amd._abi import pack_observable
from openqarp_amd.simulator import DeviceState

pytestmark = pytest.mark.gpu


def _both(bra, ket, obs):
    lib = _lib.load()
    po = pack_observable(obs)
    a = (C.c_double * 2)()
    b = (C.c_double * 2)()
    _lib.check(lib.qhip_transition(bra.handle, ket.handle, po.ref, a))
    _lib.check(lib.qhip_transition_noalloc(bra.handle, ket.handle, po.ref, b))
    return complex(a[0], a[1]), complex(b[0], b[1])


def _state(n, seed, precision="complex128"):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    return DeviceState(n, precision).upload(v / np.linalg.norm(v))


@pytest.mark.parametrize("precision", ["complex128", "complex64"])
def test_noalloc_equals_transition_on_two_calls(precision):
    # Call 1: 27 qubits = 32768 chunks, two slices of the fixed partials area;
    # Hermitian path (bra is ket, real coefficients).
    print(f"noalloc {precision}: 27 qubits, 300 strings", flush=True)
    big = _state(27, 1, precision)
    obs = W.random_pauli_sum(27, 300, 2)
    ref, got = _both(big, big, obs)
    assert got == ref
    # Call 2: a smaller register, distinct bra and ket, complex coefficients
    # (the general path); the array keeps call 1's contents, which must not leak.
    print(f"noalloc {precision}: 14 qubits, transition", flush=True)
    bra, ket = _state(14, 3, precision), _state(14, 4, precision)
    obs2 = [(p, c * complex(1.0, -0.4)) for p, c in W.random_pauli_sum(14, 80, 5)]
    ref2, got2 = _both(bra, ket, obs2)
    assert got2 == ref2
    ref3, got3 = _both(ket, ket, [([(0, "Z")], 0.5)])
    assert got3 == ref3


def test_noalloc_rejects_observables_beyond_capacity():
    st = _state(13, 6)
    # 4100 distinct flip masks: X on the set bits of 1..4100.
    obs = [([(q, "X") for q in range(13) if (m >> q) & 1], 1.0 / m) for m in range(1, 4101)]
    lib = _lib.load()
    po = pack_observable(obs)
    out = (C.c_double * 2)()
    with pytest.raises(ValueError, match="at most 4096 groups"):
        _lib.check(lib.qhip_transition_noalloc(st.handle, st.handle, po.ref, out))
    ref, got = _both(st, st, obs[:200])
    assert got == ref
