"""Shot-batched trajectories: QarpSimulator::run for circuits with mid-circuit
Measure, Reset, classical conditions and branches
(src/qarpx/src/simulator/qarp_simulator.cpp:1038-1211; per command,
apply_command_trajectory :600-689; branches, build_branch_table :298-350).

The unitary prefix before the first trajectory command runs once.  Shots then
run in batches of B = 2^b inside one (n + b)-qubit device state, shot j in
the amplitude block [j·2^n, (j+1)·2^n):

- a run of unconditional gates is one fused apply on the whole state (its
  qubits are < n, so every shot advances together); when only some shots
  execute it (inside a branch, or a conditional gate), the others are
  restored from a copy afterwards (qhip_shots_restore);
- Measure / Reset: the probability of |0> per shot on the device
  (qhip_shots_prob0), one uniform per executing shot from its own stream
  std::mt19937(base_s + shot) (qhip_mt_uniforms), outcome = u >= p0, then
  the per-shot projection and renormalisation on the device
  (qhip_shots_project);
- the final draw per shot is the reference's inline CDF walk, one device
  thread per shot (qhip_shots_sample): probabilities |a|^2 = re*re + im*im,
  their running sum in index order, threshold = u · total, first index whose
  running sum reaches it.

With an active noise model (noise.py) the per-shot loop starts at the first
command, as the reference's does (split 0, qarp_simulator.cpp:1045), and every
applied gate whose type carries a channel is followed by that channel on the
shots that applied it (apply_post_gate_noise, :697-722): each shot draws from
its own stream, and the sampled Pauli or Kraus operator is applied per shot
(qhip_shots_apply1q, with qhip_shots_rho1 for the Kraus probabilities).

Random numbers are consumed per shot exactly as the reference does (one per
executed Measure or Reset, one per noisy gate — plus a uniform integer when a
many-qubit Pauli channel fires — and one for the final draw), so outcomes
agree with the reference's except where a comparison lands within rounding of
a probability."""
from __future__ import annotations

import ctypes as C
import secrets
from typing import Optional, Sequence

import numpy as np

from . import _lib
from .ir import Command, GateType, SamplingResult, cbit_register_width
from .program import needs_trajectory

MAX_BATCH_AMPS = 1 << 24        # amplitudes of one batched state


def _parse(cmds: Sequence[Command], i: int = 0, stop=(), noise=None) -> tuple:
    """Items of cmds[i:] up to a stop marker: ("gates", [..]) runs,
    ("op", cmd, channel) for Measure / Reset / conditional commands and for
    gates followed by a noise channel (channel None otherwise), and
    ("branch", begin, then_items, else_items)."""
    items = []
    run: list = []

    def flush():
        if run:
            items.append(("gates", list(run)))
            run.clear()
    while i < len(cmds):
        c = cmds[i]
        channel = noise.channel_for(c) if noise is not None else None
        if c.gate == GateType.BranchBegin:
            flush()
            then, i = _parse(cmds, i + 1, (GateType.BranchElse, GateType.BranchEnd), noise)
            els = []
            if cmds[i].gate == GateType.BranchElse:
                els, i = _parse(cmds, i + 1, (GateType.BranchEnd,), noise)
            items.append(("branch", c, then, els))
            i += 1
        elif c.gate in (GateType.BranchElse, GateType.BranchEnd):
            if c.gate in stop:
                flush()
                return items, i
            raise RuntimeError(f"branch table: {c.gate.name} without matching BranchBegin")
        elif c.gate in (GateType.Measure, GateType.Reset) or c.condition_bits or channel is not None:
            flush()
            items.append(("op", c, channel))
            i += 1
        elif c.gate == GateType.Barrier:
            i += 1
        else:
            run.append(c)
            i += 1
    if stop:
        raise RuntimeError("branch table: BranchBegin without matching BranchEnd")
    flush()
    return items, i


def _draws(items) -> int:
    """Uniforms a shot can draw along every branch (noiseless items)."""
    n = 0
    for it in items:
        if it[0] == "op" and it[1].gate in (GateType.Measure, GateType.Reset):
            n += 1
        elif it[0] == "branch":
            n += _draws(it[2]) + _draws(it[3])
    return n


class _Precomputed:
    """The uniforms of noiseless shots, drawn ahead on the library's host side
    (qhip_mt_uniforms), read per shot in order."""

    def __init__(self, u: np.ndarray):
        self.u = u
        self.ptr = np.zeros(u.shape[0], dtype=np.int64)

    def uniform(self, idx: np.ndarray) -> np.ndarray:
        out = self.u[idx, self.ptr[idx]]
        self.ptr[idx] += 1
        return out


_X = np.array([0, 0, 1, 0, 1, 0, 0, 0], dtype=np.float64)
_Y = np.array([0, 0, 0, -1, 0, 1, 0, 0], dtype=np.float64)
_Z = np.array([1, 0, 0, 0, 0, 0, -1, 0], dtype=np.float64)
_PAULI_MATS = {1: _X, 2: _Y, 3: _Z}


def _holds(bits, values, regs: np.ndarray) -> np.ndarray:
    ok = np.ones(regs.shape[0], dtype=bool)
    for b, v in zip(bits, values):
        ok &= regs[:, b] == bool(v)
    return ok


class _Batch:
    def __init__(self, sim, n: int, b: int, n_cbits: int, streams, plans: dict):
        from .simulator import DeviceState
        self.sim = sim
        self.plans = plans   # per call: id(run) -> (run, CompiledCircuit on n + b qubits), bare conditionals
        self.n, self.stride, self.B = n, 1 << n, 1 << b
        self.state = DeviceState(n + b, sim.precision)
        self.saved: Optional[DeviceState] = None
        self.regs = np.zeros((self.B, n_cbits), dtype=bool)
        self.streams = streams                  # per-shot uniforms: _Precomputed or ShotStreams
        self.lib = _lib.load()

    def apply(self, cmds) -> None:
        """One fused, scheduled apply of a command run, compiled once per run
        for every batch (all batches share the register width)."""
        from .simulator import CompiledCircuit
        plan = self.plans.get(id(cmds))
        if plan is None:
            plan = self.plans[id(cmds)] = (cmds, CompiledCircuit(self.sim, cmds, self.state.n_qubits))
        plan[1].apply(self.state)

    def bare(self, item) -> list:
        """The command of a conditional item without its condition, as a run
        (one list per item for the whole call, so its plan is compiled once)."""
        key = ("bare", id(item))
        if key not in self.plans:
            c = item[1]
            self.plans[key] = (item, [Command(c.gate, qubits=c.qubits, params=c.params, unitary=c.unitary)])
        return self.plans[key][1]

    def gates(self, cmds, active: np.ndarray) -> None:
        if not active.any():
            return
        if active.all():
            self.apply(cmds)
            return
        if self.saved is None:
            from .simulator import DeviceState
            self.saved = DeviceState(self.state.n_qubits, self.sim.precision)
        self.saved.copy_from(self.state)
        self.apply(cmds)
        act = np.ascontiguousarray(active, dtype=np.uint8)
        _lib.check(self.lib.qhip_shots_restore(self.state.handle, self.saved.handle, act.ctypes.data, self.stride))

    def collapse(self, q: int, cbits, reset: bool, active: np.ndarray) -> None:
        if not active.any():
            return
        p0 = np.zeros(self.B)
        _lib.check(self.lib.qhip_shots_prob0(self.state.handle, q, self.stride, p0.ctypes.data))
        idx = np.nonzero(active)[0]
        u = self.streams.uniform(idx)
        one = u >= p0[idx]
        with np.errstate(divide="ignore"):
            p = np.where(one, np.maximum(1.0 - p0[idx], 0.0), np.maximum(p0[idx], 0.0))
            f = np.sqrt(1.0 / p)
        action = np.full(self.B, -1, dtype=np.int8)
        action[idx] = np.where(one, 2 if reset else 1, 0)
        scale = np.ones(self.B)
        scale[idx] = f
        _lib.check(self.lib.qhip_shots_project(self.state.handle, q, self.stride, action.ctypes.data,
                                               scale.ctypes.data))
        if cbits and not reset:
            self.regs[idx, cbits[0]] = one

    def paulis(self, q: int, codes: np.ndarray) -> None:
        """Per shot the Pauli codes[j] (0 I, 1 X, 2 Y, 3 Z) on qubit q."""
        act = np.ascontiguousarray(codes != 0, dtype=np.uint8)
        if not act.any():
            return
        mats = np.zeros((self.B, 8))
        for c, m in _PAULI_MATS.items():
            mats[codes == c] = m
        _lib.check(self.lib.qhip_shots_apply1q(self.state.handle, int(q), self.stride, mats.ctypes.data,
                                               act.ctypes.data))

    def channel(self, cmd, ch, active: np.ndarray) -> None:
        """apply_channel (device/noise_model.cpp:86-186) on the shots in
        active, each drawing from its own stream."""
        from .noise import KrausChannel, PauliChannel, PauliChannel2q, PauliChannelKq
        idx = np.nonzero(active)[0]
        if idx.size == 0 or ch is None:
            return
        qs = list(cmd.qubits)
        if isinstance(ch, PauliChannel):
            if len(qs) != 1:
                raise RuntimeError("apply_channel: PauliChannel requires exactly 1 target qubit")
            r = self.streams.uniform(idx)
            a1 = ch.p_x
            a2 = a1 + ch.p_y
            a3 = a2 + ch.p_z
            codes = np.zeros(self.B, dtype=np.int64)
            codes[idx] = np.where(r < a1, 1, np.where(r < a2, 2, np.where(r < a3, 3, 0)))
            self.paulis(qs[0], codes)
        elif isinstance(ch, PauliChannel2q):
            if len(qs) != 2:
                raise RuntimeError("apply_channel: PauliChannel2q requires exactly 2 target qubits")
            r = self.streams.uniform(idx)
            acc = np.empty(15)
            t = 0.0
            for k in range(15):
                t += ch.p[k]
                acc[k] = t
            hit = r[:, None] < acc[None, :]
            first = np.where(hit.any(axis=1), hit.argmax(axis=1), -1)
            packed = np.zeros(self.B, dtype=np.int64)
            packed[idx] = np.where(first >= 0, first + 1, 0)
            self.paulis(qs[0], packed & 3)
            self.paulis(qs[1], (packed >> 2) & 3)
        elif isinstance(ch, PauliChannelKq):
            k = len(qs)
            if k == 0 or k > 31:
                return
            r = self.streams.uniform(idx)
            codes = np.zeros((self.B, k), dtype=np.int64)
            for j in idx[r < ch.p_total]:
                code = self.streams.uniform_int(int(j), 1, (1 << (2 * k)) - 1)
                codes[j] = [(code >> (2 * i)) & 3 for i in range(k)]
            for i in range(k):
                self.paulis(qs[i], codes[:, i])
        elif isinstance(ch, KrausChannel):
            if len(qs) != 1:
                raise RuntimeError("apply_channel: KrausChannel requires exactly 1 target qubit")
            if not ch.K:
                return
            rho4 = np.zeros((self.B, 4))
            _lib.check(self.lib.qhip_shots_rho1(self.state.handle, qs[0], self.stride, rho4.ctypes.data))
            mats = np.zeros((self.B, 8))
            act = np.zeros(self.B, dtype=np.uint8)
            for j in idx:
                rho = np.array([[rho4[j, 0], complex(rho4[j, 2], rho4[j, 3])],
                                [complex(rho4[j, 2], -rho4[j, 3]), rho4[j, 1]]], dtype=np.complex128)
                probs = [max(0.0, float(np.trace(K @ rho @ K.conj().T).real)) for K in ch.K]
                total = 0.0
                for p in probs:
                    total += p
                if total <= 0.0:
                    continue                    # no draw, as the reference returns before its uniform
                probs = [p / total for p in probs]
                u = float(self.streams.uniform(np.array([j]))[0])
                acc = 0.0
                for i, p in enumerate(probs):
                    acc += p
                    if u < acc:
                        p_unnorm = p * total
                        if p_unnorm > 0.0:
                            m = np.asarray(ch.K[i], dtype=np.complex128) * (1.0 / np.sqrt(p_unnorm))
                            mats[j] = m.view(np.float64).ravel()
                            act[j] = 1
                        break
            if act.any():
                _lib.check(self.lib.qhip_shots_apply1q(self.state.handle, qs[0], self.stride, mats.ctypes.data,
                                                       act.ctypes.data))
        else:
            raise RuntimeError(f"apply_channel: unknown channel {ch!r}")

    def run(self, items, active: np.ndarray) -> None:
        for it in items:
            if it[0] == "gates":
                self.gates(it[1], active)
            elif it[0] == "branch":
                hold = _holds(it[1].condition_bits, it[1].condition_values, self.regs)
                self.run(it[2], active & hold)
                if it[3]:
                    self.run(it[3], active & ~hold)
            else:
                c = it[1]
                act = active & _holds(c.condition_bits, c.condition_values, self.regs) if c.condition_bits \
                    else active
                if c.gate in (GateType.Measure, GateType.Reset):
                    if not c.qubits:
                        raise RuntimeError(f"trajectory: {c.gate.name} requires a qubit")
                    self.collapse(c.qubits[0], list(c.cbits), c.gate == GateType.Reset, act)
                else:
                    self.gates(self.bare(it), act)
                    if it[2] is not None:
                        self.channel(c, it[2], act)

    def sample(self, count: int) -> np.ndarray:
        """Each shot's final draw on the device (qhip_shots_sample), with the
        shot's next uniform."""
        u = np.ascontiguousarray(self.streams.uniform(np.arange(count)))
        out = np.zeros(count, dtype=np.uint64)
        _lib.check(self.lib.qhip_shots_sample(self.state.handle, self.stride, u.ctypes.data, count,
                                              out.ctypes.data))
        return out


def run_trajectories(sim, commands: Sequence[Command], n_qubits: int, n_shots: int,
                     seed: Optional[int] = None, initial_state=None, noise=None) -> SamplingResult:
    """QarpSimulator::run's trajectory path on the device (shots in batches);
    ``noise`` an active NoiseModel or None."""
    from .noise import ShotStreams
    n = int(n_qubits)
    n_shots = int(n_shots)
    if n_shots < 0:
        raise ValueError("QarpSimulator::run: n_shots must be non-negative")
    base_s = (secrets.randbits(32) if seed is None else int(seed)) & 0xFFFFFFFF
    commands = list(commands)
    if noise is not None and not noise.active():
        noise = None
    split = 0 if noise is not None else next((i for i, c in enumerate(commands) if needs_trajectory(c)),
                                             len(commands))
    items, _ = _parse(commands[split:], noise=noise)
    n_draws = _draws(items) + 1
    n_cbits = cbit_register_width(commands)
    base = sim.new_state(n, initial_state, "run")
    sim.apply(commands[:split], base)
    res = SamplingResult(n_qubits=n, n_shots=n_shots)
    res.n_cbits = n_cbits
    counts: dict = {}
    history: list = []
    lib = _lib.load()
    plans: dict = {}
    # Shots per batch: enough for all of them, within MAX_BATCH_AMPS amplitudes.
    b = min(max(n_shots - 1, 0).bit_length(), max(0, MAX_BATCH_AMPS.bit_length() - 1 - n))
    for shot0 in range(0, n_shots, 1 << b):
        count = min(1 << b, n_shots - shot0)
        if noise is not None:
            streams = ShotStreams(base_s, shot0, 1 << b, 2 * n_draws + 64)
        else:
            u = np.zeros((1 << b, n_draws))
            for s in range(count):
                _lib.check(lib.qhip_mt_uniforms(C.c_uint32((base_s + shot0 + s) & 0xFFFFFFFF), n_draws,
                                                u[s].ctypes.data))
            streams = _Precomputed(u)
        batch = _Batch(sim, n, b, n_cbits, streams, plans)
        _lib.check(lib.qhip_shots_broadcast(batch.state.handle, base.handle))
        active = np.zeros(1 << b, dtype=bool)
        active[:count] = True
        batch.run(items, active)
        outcomes = batch.sample(count)
        vals, cnt = np.unique(outcomes, return_counts=True)
        for v, c in zip(vals.tolist(), cnt.tolist()):
            counts[int(v)] = counts.get(int(v), 0) + int(c)
        history.extend([list(map(bool, r)) for r in batch.regs[:count]])
    res.counts = counts
    res.cbit_history = history
    return res
