"""Readout index work, bit for bit: the EXACT distribution keeps exactly the
indices numpy's ``abs(sv)**2 > 1e-12`` keeps and returns the same
probabilities (src/qarp/engines/_engine.py:100-102), and column compaction
keeps exactly the entries with ``hypot(re, im) > tol``
(transition_columns.cu:268-276) — including amplitudes placed one ulp on
either side of the thresholds."""
import ctypes as C

import numpy as np
import pytest

from openqarp_amd import _lib
from openqarp_amd.simulator import DeviceState, QarpSimulator

pytestmark = pytest.mark.gpu


def _wide_magnitudes(n, seed):
    """Pairs (re, im) over many binades and ratios, as amplitudes."""
    rng = np.random.default_rng(seed)
    m = 1 << n
    ex = rng.uniform(-25, 0, m)
    re = rng.uniform(1, 2, m) * 2.0 ** np.floor(ex) * rng.choice([-1, 1], m)
    kind = rng.integers(0, 3, m)
    ratio = np.where(kind == 0, rng.uniform(0, 1, m),
                     np.where(kind == 1, 1 + 2.0 ** rng.uniform(-40, -1, m), 2.0 ** rng.uniform(-80, -20, m)))
    im = re * ratio * rng.choice([-1, 1], m)
    swap = rng.random(m) < 0.5
    re, im = np.where(swap, im, re), np.where(swap, re, im)
    return re + 1j * im


def _edge_amplitudes(thr):
    """Amplitudes whose numpy |a|^2 lies within a few ulps of thr, with zero and
    nonzero imaginary parts."""
    r = np.sqrt(thr)
    out = []
    for k in range(-6, 7):
        x = r
        for _ in range(abs(k)):
            x = np.nextafter(x, np.inf if k > 0 else 0.0)
        out.append(complex(x, 0.0))
        out.append(complex(x * np.sqrt(0.5), x * np.sqrt(0.5)))
        out.append(complex(x * 0.6, x * 0.8))
        out.append(complex(0.0, -x))
    return np.array(out)


@pytest.mark.parametrize("precision", ["complex128", "complex64"])
def test_exact_readout_is_numpys_abs_squared(precision):
    n = 20
    amps = _wide_magnitudes(n, 3)
    edge = _edge_amplitudes(1e-12)
    amps[: edge.size] = edge
    if precision == "complex64":
        amps = amps.astype(np.complex64).astype(np.complex128)
    st = DeviceState(n, precision).upload(amps)
    sim = QarpSimulator(precision=precision)
    for thr in (1e-12, 0.0, 3.7e-9):
        keys, probs = sim.exact_state(st, thr)
        want = np.abs(amps) ** 2
        (idx,) = np.nonzero(want > thr)
        assert np.array_equal(keys, idx)
        assert np.array_equal(probs, want[idx])


def test_compaction_is_hypot_above_the_tolerance():
    n = 18
    amps = _wide_magnitudes(n, 5)
    tol = 1e-6
    edge = _edge_amplitudes(tol * tol)
    amps[: edge.size] = edge
    st = DeviceState(n).upload(amps)
    lib = _lib.load()
    cnt = C.c_int64()
    mp = C.c_double()
    _lib.check(lib.qhip_compact(st.handle, amps.size, tol, None, None, 0, C.byref(cnt), C.byref(mp)))
    keys = np.zeros(cnt.value, np.int64)
    vals = np.zeros(cnt.value, np.complex128)
    _lib.check(lib.qhip_compact(st.handle, amps.size, tol, keys.ctypes.data, vals.ctypes.data, cnt.value,
                                C.byref(cnt), None))
    mag = np.hypot(amps.real, amps.imag)
    (idx,) = np.nonzero(mag > tol)
    assert np.array_equal(keys, idx)
    assert np.array_equal(vals, amps[idx])
    assert mp.value == mag[mag <= tol].max()
