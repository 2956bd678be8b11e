"""The streaming tile kernel (persistent workgroups that load their next
tile while the current tile's ops run, matrix-core fragments staged in LDS)
against the one-tile-per-workgroup kernel and the oracle.  Both kernels run
the same arithmetic in the same order per amplitude, so their states are
bit-equal; the oracle bounds them at 1e-12 (complex128) / 2e-5 (complex64).
States of 2^20 .. 2^22 amplitudes give every persistent workgroup several
tiles."""
import numpy as np
import pytest

import _oracle as O
from openqarp_amd import workloads as W
from openqarp_amd._abi import pack_stream
from openqarp_amd.simulator import CompiledCircuit, QarpSimulator

pytestmark = pytest.mark.gpu


def _run(cmds, n, precision, tile, k, stream, low=0):
    sim = QarpSimulator(precision=precision, tile_qubits=tile, low_qubits=low, stream_tiles=stream)
    sim.fusion_max_qubits, sim.fusion_min_qubits = k, 0
    out = sim.statevector(cmds, n)
    return out, sim.last_stats


@pytest.mark.parametrize("precision,tile", [("complex128", 10), ("complex128", 11), ("complex128", 12),
                                            ("complex64", 11), ("complex64", 12), ("complex64", 13)])
@pytest.mark.parametrize("k", [2, 3, 4, 5])
def test_streaming_kernel_is_bit_equal_to_one_tile_kernel(precision, tile, k):
    n = 21 if precision == "complex128" else 22
    cmds = W.brickwork_haar(n, 8, 100 + k + tile)
    on, st_on = _run(cmds, n, precision, tile, k, 1)
    off, st_off = _run(cmds, n, precision, tile, k, 0)
    assert st_off["stream_passes"] == 0
    # Passes without matrix-core ops stay on the one-tile kernel below the
    # widest tile.
    assert (st_on["stream_passes"] > 0) == (k >= 3), st_on
    assert st_on["passes"] == st_off["passes"]
    assert np.array_equal(on, off)


@pytest.mark.parametrize("precision,tile,k", [("complex128", 12, 4), ("complex128", 11, 5), ("complex64", 13, 4),
                                              ("complex128", 13, 2), ("complex128", 13, 4), ("complex128", 13, 5),
                                              ("complex64", 14, 3), ("complex64", 14, 4)])
def test_streaming_kernel_matches_oracle(precision, tile, k):
    """Including the widest tiles (2^13 complex128 / 2^14 complex64
    amplitudes, 128 KiB), which only the streaming kernel runs; k = 5 passes
    there hold more matrix-core fragments than the LDS staging area, so some
    ops read theirs from global memory."""
    n = 20
    cmds = W.brickwork_haar(n, 10, 7 + k) + W.brickwork_haar(n, 4, 70 + k)
    got, st = _run(cmds, n, precision, tile, k, 1, low=3)
    assert st["stream_passes"] > 0
    if tile >= 13 + (precision == "complex64"):
        assert st["stream_passes"] == st["passes"]
    ref = O.statevector(pack_stream(cmds), n, k, 0)
    assert np.abs(got - ref).max() < (1e-12 if precision == "complex128" else 2e-5)


def test_widest_tile_needs_the_streaming_kernel():
    cmds = W.brickwork_haar(16, 2, 1)
    with pytest.raises(Exception, match="streaming tile kernel"):
        _run(cmds, 16, "complex128", 13, 2, 0)
    got, st = _run(cmds, 16, "complex128", 13, 2, -1)   # the default streams the widest tile only
    assert st["stream_passes"] == st["passes"] > 0
    assert np.abs(got - O.statevector(pack_stream(cmds), 16, 2, 0)).max() < 1e-12
    _, st = _run(cmds, 16, "complex128", 12, 4, -1)
    assert st["stream_passes"] == 0


def test_streaming_kernel_in_a_captured_graph():
    """Replays of the captured schedule equal the eager apply, with eager work
    and synchronisations between the replays."""
    n, k = 21, 4
    cmds = W.brickwork_haar(n, 8, 3)
    sim = QarpSimulator(tile_qubits=12, low_qubits=3, stream_tiles=1)
    cc = CompiledCircuit(sim, cmds, n, fusion_max_qubits=k)
    st = sim.new_state(n)
    stats = cc.apply(st)
    assert stats["stream_passes"] > 0
    want = st.download()
    for _ in range(3):
        st.reset()
        cc.apply_graph(st)
        assert np.array_equal(st.download(), want)
        st.reset()
        cc.apply(st)
        assert np.array_equal(st.download(), want)
