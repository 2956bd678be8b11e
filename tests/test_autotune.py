"""The autotuner: its candidate grid on the host, and on the device that the
pick is the fastest measured trial of that grid (ties to the first in grid
order), is cached per circuit, and keeps parity with the oracle when the
picked configuration runs."""
import numpy as np
import pytest

import _oracle as O
from openqarp_amd import workloads as W
from openqarp_amd._abi import pack_stream
from openqarp_amd.autotune import candidates, default_lows, fingerprint


def test_candidate_grid():
    c = candidates(30, "complex128")
    assert c == [(k, T, low) for k in (2, 3, 4, 5) for T in (11, 12) for low in (3, 4)]
    c64 = candidates(34, "complex64", widths=(2, 4))
    assert c64 == [(k, T, low) for k in (2, 4) for T in (12, 13) for low in (4, 5)]
    assert default_lows("complex64") == (4, 5)
    # Tiles wider than the register are skipped; with none left, one default.
    assert candidates(11, "complex128", widths=(3,)) == [(3, 11, 3), (3, 11, 4)]
    assert candidates(6, "complex128", widths=(3,)) == [(3, 6, 4)]
    assert candidates(20, "complex128", widths=(4,), tiles=(9,), lows=(2,)) == [(4, 9, 2)]


def test_fingerprint_follows_structure_not_angles():
    a = W.brickwork_haar(8, 3, 1)
    b = W.brickwork_haar(8, 3, 1)
    assert fingerprint(a) == fingerprint(b)
    c = W.brickwork_haar(8, 3, 2)
    assert fingerprint(a) != fingerprint(c)


@pytest.mark.gpu
def test_tune_picks_the_fastest_trial_and_keeps_parity():
    from openqarp_amd.autotune import _cache, tune
    from openqarp_amd.simulator import CompiledCircuit, QarpSimulator
    n = 16
    cmds = W.brickwork_haar(n, 6, 7)
    sim = QarpSimulator()
    st = sim.new_state(n)
    out = tune(sim, cmds, st, widths=(2, 3, 4), tiles=(10, 11), lows=(3, 4))
    trials = out["trials"]
    grid = candidates(n, "complex128", (2, 3, 4), (10, 11), (3, 4))
    assert [(t["fusion_max_qubits"], t["tile_qubits"], t["low_qubits"]) for t in trials[:len(grid)]] == grid
    assert all(t["stream_tiles"] == -1 for t in trials[:len(grid)])
    # One more trial: the grid's winner with its matrix-core passes streamed.
    grid_best = min(trials[:len(grid)], key=lambda t: t["ms"])
    if grid_best["mfma_ops"] > 0:
        extra = trials[len(grid):]
        assert len(extra) == 1 and extra[0]["stream_tiles"] == 1 and extra[0]["stream_passes"] > 0
        assert [extra[0][key] for key in ("fusion_max_qubits", "tile_qubits", "low_qubits")] == \
            [grid_best[key] for key in ("fusion_max_qubits", "tile_qubits", "low_qubits")]
    else:
        assert len(trials) == len(grid)
    best = out["best"]
    fastest = min(t["ms"] for t in trials)
    assert best["ms"] == fastest
    assert best is next(t for t in trials if t["ms"] == fastest)
    assert all(t["passes"] >= 1 and t["ms"] > 0 for t in trials)
    assert tune(sim, cmds, st, widths=(2, 3, 4), tiles=(10, 11), lows=(3, 4)) is out   # cached
    assert len(_cache) >= 1
    cc = CompiledCircuit(sim, cmds, n, fusion_max_qubits=best["fusion_max_qubits"],
                         tile_qubits=best["tile_qubits"], low_qubits=best["low_qubits"],
                         stream_tiles=best["stream_tiles"])
    st.reset()
    cc.apply(st)
    ref = O.statevector(pack_stream(cmds), n, best["fusion_max_qubits"], 0)
    assert np.abs(st.download() - ref).max() < 1e-12
