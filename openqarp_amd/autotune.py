"""Picks fusion width and tile shape for a circuit by timing candidate
schedules on the device.

Each candidate (fusion width k, tile qubits T, contiguous low qubits) is
compiled once and applied once on the caller's state with HIP-event timing
of every pass; the fastest total wins.  The winner is then timed once more
with its matrix-core passes on the streaming tile kernel (stream_tiles = 1),
which is kept when faster.  Results are cached per (register
width, precision, circuit fingerprint).  Applying a candidate changes the
state's amplitudes, not the work done, so the caller resets the state
afterwards when its contents matter.
"""
from __future__ import annotations

import hashlib
from typing import Iterable, Optional, Sequence

import numpy as np

from .ir import Command
from .simulator import CompiledCircuit, DeviceState, QarpSimulator

_cache: dict = {}


def fingerprint(commands: Sequence[Command]) -> str:
    h = hashlib.sha1()
    for c in commands:
        h.update(bytes([int(c.gate)]))
        h.update(np.asarray(c.qubits, np.int64).tobytes())
    return h.hexdigest()


def default_lows(precision: str) -> tuple:
    """Contiguous low qubits tried: runs of 128 B and 256 B per tile row
    (8 or 16 complex128 amplitudes, 16 or 32 complex64)."""
    return (3, 4) if precision == "complex128" else (4, 5)


def candidates(n: int, precision: str, widths: Iterable[int] = (2, 3, 4, 5),
               tiles: Optional[Iterable[int]] = None, lows: Optional[Iterable[int]] = None) -> list:
    base_t = 12 if precision == "complex128" else 13
    tiles = tuple(tiles) if tiles else (base_t - 1, base_t)
    lows = tuple(lows) if lows else default_lows(precision)
    out = []
    for k in widths:
        for T in tiles:
            for low in lows:
                if T <= n:
                    out.append((k, T, low))
    return out or [(3, min(n, base_t), 4)]


def tune(sim: QarpSimulator, commands: Sequence[Command], state: DeviceState,
         widths: Iterable[int] = (2, 3, 4, 5), repeats: int = 1, tiles: Optional[Iterable[int]] = None,
         lows: Optional[Iterable[int]] = None, try_stream: bool = True) -> dict:
    widths = tuple(widths)
    lows = tuple(lows) if lows else default_lows(sim.precision)
    tiles = tuple(tiles) if tiles else None
    key = (state.n_qubits, sim.precision, fingerprint(commands), widths, tiles, lows, try_stream)
    if key in _cache:
        return _cache[key]
    results = []

    def trial(k, T, low, stream):
        cc = CompiledCircuit(sim, commands, state.n_qubits, fusion_max_qubits=k, tile_qubits=T,
                             low_qubits=low, stream_tiles=stream)
        cc.apply(state)                         # warm (first launch of each kernel)
        ms = min(float(cc.apply_timed(state).sum()) for _ in range(repeats))
        results.append({"fusion_max_qubits": k, "tile_qubits": T, "low_qubits": low, "stream_tiles": stream,
                        "ms": ms, "passes": cc.stats["passes"], "fused_ops": cc.stats["fused_ops"],
                        "mfma_ops": cc.stats["mfma_ops"], "stream_passes": cc.stats["stream_passes"]})

    for k, T, low in candidates(state.n_qubits, sim.precision, widths, tiles, lows):
        trial(k, T, low, sim.stream_tiles)
    best = min(results, key=lambda r: r["ms"])
    if try_stream and sim.stream_tiles != 0 and best["mfma_ops"] > 0:
        trial(best["fusion_max_qubits"], best["tile_qubits"], best["low_qubits"], 1)
        best = min(results, key=lambda r: r["ms"])
    out = {"best": best, "trials": results}
    _cache[key] = out
    return out
