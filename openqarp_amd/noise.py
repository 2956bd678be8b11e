"""Noise channels for the trajectory path, and the per-shot random streams.

The model mirrors the reference's (src/qarpx/include/qarpx/device/
noise_model.h): a channel per gate type, looked up after every applied gate
except Measure, Reset, Barrier, GPhase and the branch markers
(apply_post_gate_noise, qarp_simulator.cpp:697-722); ``enabled`` switches it
off; ``+`` composes (the right operand wins per gate, ``enabled`` is ANDed);
the builders are make_depolarizing, make_pauli, make_bit_flip and
make_amplitude_damping (device/noise_model.cpp:221-299).  Each channel draws
from the shot's own stream as apply_channel does (:86-186):

- PauliChannel (one qubit): one uniform r; X if r < p_x, Y if r < p_x + p_y,
  Z if r < p_x + p_y + p_z, else nothing;
- PauliChannel2q: one uniform r against the running sum of the 15
  probabilities, string id + 1 = p0 | p1 << 2 (p0 on the first qubit);
- PauliChannelKq (variadic MCZ under make_depolarizing): one uniform r, fires
  when r < p_total, then a uniform integer in [1, 4^k - 1] picks the string
  two bits per qubit;
- KrausChannel (one qubit): the reduced density matrix of the qubit gives
  p_i = max(0, Re tr(K_i rho K_i^dag)); normalised by their sum, one uniform
  picks K_i, applied as K_i / sqrt(p_i).

``ShotStreams`` gives every shot of a run its own std::mt19937(base + shot)
word stream (numpy's MT19937 bit generator set to the C++ standard's seeding
state), uniforms as libstdc++'s generate_canonical<double, 53> forms them
from two words, and uniform integers as libstdc++'s uniform_int_distribution
draws them (Lemire's multiply-and-reject on 32-bit words, upscaling for
ranges of 2^32 and more) — the streams the reference's shots consume.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Callable, Dict, Sequence, Union

import numpy as np

from .ir import Command, GateType

# ── channels ────────────────────────────────────────────────────────────────


@dataclass(frozen=True)
class PauliChannel:
    p_x: float = 0.0
    p_y: float = 0.0
    p_z: float = 0.0


@dataclass(frozen=True)
class PauliChannel2q:
    p: tuple = (0.0,) * 15       # id = (p0 | p1 << 2) - 1, letters I, X, Y, Z = 0..3


@dataclass(frozen=True)
class PauliChannelKq:
    p_total: float = 0.0


@dataclass(frozen=True)
class KrausChannel:
    K: tuple = ()                # 2x2 complex128 arrays


Channel = Union[None, PauliChannel, PauliChannel2q, PauliChannelKq, KrausChannel]
NoiseChannel = Callable[[Command], Channel]

_NOISELESS = {GateType.Measure, GateType.Reset, GateType.Barrier, GateType.GPhase, GateType.BranchBegin,
              GateType.BranchElse, GateType.BranchEnd}

_TWO = {GateType.CX, GateType.CY, GateType.CZ, GateType.SWAP, GateType.ECR, GateType.iSWAP, GateType.iSWAPdg,
        GateType.CH, GateType.CS, GateType.CSdg, GateType.CSX, GateType.CSXdg, GateType.CRx, GateType.CRy,
        GateType.CRz, GateType.CP, GateType.RZZ, GateType.RXX, GateType.RYY, GateType.CU, GateType.MCZ}
_ZERO = {GateType.GPhase, GateType.BranchBegin, GateType.BranchElse, GateType.BranchEnd}


def gate_num_qubits(g: GateType) -> int:
    """Minimum operand count of a gate type (core/gates.h:130-153): 2 for the
    two-qubit gates and MCZ, 3 for CCX and CSWAP, 0 for GPhase and the branch
    markers, 1 otherwise."""
    g = GateType(g)
    if g in _TWO:
        return 2
    if g in (GateType.CCX, GateType.CSWAP):
        return 3
    return 0 if g in _ZERO else 1


class NoiseModel:
    """A channel per gate type (a constant or a function of the command)."""

    def __init__(self):
        self.per_gate: Dict[GateType, NoiseChannel] = {}
        self.enabled = True

    def set_channel(self, gate: GateType, channel) -> None:
        self.per_gate[GateType(gate)] = channel if callable(channel) else (lambda cmd, c=channel: c)

    def has_channel(self, gate: GateType) -> bool:
        return GateType(gate) in self.per_gate

    def has_any_channel(self) -> bool:
        return bool(self.per_gate)

    def active(self) -> bool:
        return bool(self.enabled) and self.has_any_channel()

    def channel_for(self, cmd: Command) -> Channel:
        """The channel applied after ``cmd``, or None."""
        if not self.enabled or cmd.gate in _NOISELESS:
            return None
        fn = self.per_gate.get(cmd.gate)
        return None if fn is None else fn(cmd)

    def __iadd__(self, other: "NoiseModel") -> "NoiseModel":
        self.per_gate.update(other.per_gate)
        self.enabled = self.enabled and other.enabled
        return self

    def __add__(self, other: "NoiseModel") -> "NoiseModel":
        out = NoiseModel()
        out.per_gate = dict(self.per_gate)
        out.enabled = self.enabled
        out += other
        return out


def make_depolarizing(p: float, gates: Sequence[GateType]) -> NoiseModel:
    if not (0.0 <= p <= 1.0):
        raise ValueError("make_depolarizing: p must lie in [0, 1]")
    nm = NoiseModel()
    for g in gates:
        g = GateType(g)
        if g == GateType.MCZ:
            nm.set_channel(g, PauliChannelKq(p))
            continue
        k = gate_num_qubits(g)
        if k == 1:
            nm.set_channel(g, PauliChannel(p / 3.0, p / 3.0, p / 3.0))
        elif k == 2:
            nm.set_channel(g, PauliChannel2q((p / 15.0,) * 15))
        else:
            raise ValueError(f"make_depolarizing: gate '{g.name}' has unsupported qubit count")
    return nm


def make_pauli(p_x: float, p_y: float, p_z: float, gates: Sequence[GateType]) -> NoiseModel:
    if p_x < 0.0 or p_y < 0.0 or p_z < 0.0 or p_x + p_y + p_z > 1.0:
        raise ValueError("make_pauli: probabilities must be non-negative and sum to ≤ 1")
    nm = NoiseModel()
    for g in gates:
        g = GateType(g)
        if gate_num_qubits(g) != 1:
            raise ValueError(f"make_pauli: '{g.name}' is not 1-qubit")
        nm.set_channel(g, PauliChannel(p_x, p_y, p_z))
    return nm


def make_bit_flip(p: float, gates: Sequence[GateType]) -> NoiseModel:
    return make_pauli(p, 0.0, 0.0, gates)


def make_amplitude_damping(p: float, gates: Sequence[GateType]) -> NoiseModel:
    if not (0.0 <= p <= 1.0):
        raise ValueError("make_amplitude_damping: p must lie in [0, 1]")
    k0 = np.array([[1.0, 0.0], [0.0, np.sqrt(1.0 - p)]], dtype=np.complex128)
    k1 = np.array([[0.0, np.sqrt(p)], [0.0, 0.0]], dtype=np.complex128)
    nm = NoiseModel()
    for g in gates:
        g = GateType(g)
        if gate_num_qubits(g) != 1:
            raise ValueError(f"make_amplitude_damping: '{g.name}' is not 1-qubit")
        nm.set_channel(g, KrausChannel((k0, k1)))
    return nm


# ── per-shot random streams ─────────────────────────────────────────────────

def _std_mt19937_keys(seeds: np.ndarray) -> np.ndarray:
    """The 624-word states std::mt19937(seed) starts in, one row per seed (the
    standard's initialisation recurrence, run for all seeds at once)."""
    seeds = np.asarray(seeds, dtype=np.uint64) & np.uint64(0xFFFFFFFF)
    keys = np.empty((seeds.size, 624), dtype=np.uint32)
    x = seeds.copy()
    keys[:, 0] = x
    mult, mask = np.uint64(1812433253), np.uint64(0xFFFFFFFF)
    for i in range(1, 624):
        x = (mult * (x ^ (x >> np.uint64(30))) + np.uint64(i)) & mask
        keys[:, i] = x
    return keys


def _generator(key: np.ndarray) -> np.random.MT19937:
    """numpy's MT19937 bit generator in a std::mt19937 state (position at the
    end, so the first draw regenerates the block as the standard's does)."""
    bg = np.random.MT19937()
    bg.state = {"bit_generator": "MT19937", "state": {"key": key, "pos": 624}}
    return bg


class ShotStreams:
    """Word streams of shots base_s + shot0 + j, j < count, read ahead in
    blocks (a shot that runs out is extended from its own generator)."""

    def __init__(self, base_s: int, shot0: int, count: int, words_per_shot: int = 64):
        self.count = count
        seeds = (int(base_s) + int(shot0) + np.arange(count, dtype=np.uint64)) & np.uint64(0xFFFFFFFF)
        self.gens = [_generator(k) for k in _std_mt19937_keys(seeds)]
        w = max(8, int(words_per_shot))
        self.words = np.stack([g.random_raw(w) for g in self.gens]) if count else np.zeros((0, w), np.uint64)
        self.pos = np.zeros(count, dtype=np.int64)

    def _ensure(self, idx: np.ndarray, need: int) -> None:
        short = idx[self.pos[idx] + need > self.words.shape[1]]
        if short.size == 0:
            return
        grow = max(self.words.shape[1], need + 8)
        extra = np.stack([self.gens[j].random_raw(grow) for j in range(self.count)])
        self.words = np.concatenate([self.words, extra], axis=1)

    def word(self, j: int) -> int:
        self._ensure(np.array([j]), 1)
        v = int(self.words[j, self.pos[j]])
        self.pos[j] += 1
        return v

    def uniform(self, idx: np.ndarray) -> np.ndarray:
        """One std::uniform_real_distribution<double>(0, 1) draw per shot in
        idx: (w0 + w1 * 2^32) / 2^64 in double, below 1."""
        idx = np.asarray(idx, dtype=np.int64)
        if idx.size == 0:
            return np.zeros(0)
        self._ensure(idx, 2)
        p = self.pos[idx]
        w0 = self.words[idx, p].astype(np.float64)
        w1 = self.words[idx, p + 1].astype(np.float64)
        self.pos[idx] += 2
        u = (w0 + w1 * 4294967296.0) / 18446744073709551616.0
        return np.where(u >= 1.0, np.nextafter(1.0, 0.0), u)

    def uniform_int(self, j: int, lo: int, hi: int) -> int:
        """std::uniform_int_distribution<uint64_t>(lo, hi) on shot j's stream."""
        return lo + self._below(j, hi - lo)

    def _below(self, j: int, r: int) -> int:
        """Uniform in [0, r] from 32-bit words."""
        if r < 0xFFFFFFFF:
            n = r + 1
            prod = self.word(j) * n
            low = prod & 0xFFFFFFFF
            if low < n:
                threshold = ((1 << 32) - n) % n
                while low < threshold:
                    prod = self.word(j) * n
                    low = prod & 0xFFFFFFFF
            return prod >> 32
        if r == 0xFFFFFFFF:
            return self.word(j)
        while True:
            tmp = (1 << 32) * self._below(j, r >> 32)
            ret = tmp + self.word(j)
            if ret <= r and ret >= tmp:
                return ret
