"""``QarpSimulator`` for the MI355X backend.

Same methods, arguments and error behaviour as the reference's
``qx.QarpSimulator`` binding (src/qarpx/python/bindings.cpp:1337-1517;
stub src/qarpx/python/qarpx.pyi:1594-1680), so ``QarpEngine``-style callers
can hold one in place of the CPU simulator.  Every method runs on the GPU
through libqarp_hip.so; states live in device memory (``DeviceState``) and
reach the host only when a method returns amplitudes.

Beyond the reference surface: ``precision="complex64"`` keeps complex64
states with FP64 reductions, and the ``*_state`` methods work on
device-resident states without host copies.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import secrets
from typing import Mapping, Optional, Sequence

import numpy as np

from . import _lib
from .program import Program, measurements_terminal, needs_trajectory
from ._abi import (QHIP_C64, QHIP_C128, PackedStream, QhipApplyStats, config_struct,
                   pack_observable, pack_stream)
from .ir import Command, GateType, SamplingResult, as_commands, cbit_register_width

Observable = Sequence[tuple]


def _env_fusion_default() -> int:
    """QARP_FUSION_MAX_QUBITS, ignored unless it parses to 0..8
    (qarp_simulator.cpp:48-63)."""
    env = os.environ.get("QARP_FUSION_MAX_QUBITS")
    if env is not None:
        try:
            k = int(env, 10)
            if 0 <= k <= QarpSimulator.MAX_FUSION_QUBITS:
                return k
        except ValueError:
            pass
    return QarpSimulator.DEFAULT_FUSION_QUBITS


class DeviceState:
    """A statevector in device memory (complex128 or complex64)."""

    __slots__ = ("_h", "n_qubits", "precision", "__weakref__")

    def __init__(self, n_qubits: int, precision: str = "complex128"):
        lib = _lib.load()
        if precision not in ("complex128", "complex64"):
            raise ValueError(f"precision must be complex128 or complex64, got {precision!r}")
        self._h = None
        h = C.c_void_p()
        _lib.check(lib.qhip_state_create(int(n_qubits), QHIP_C128 if precision == "complex128" else QHIP_C64,
                                         C.byref(h)))
        self._h = h
        self.n_qubits = int(n_qubits)
        self.precision = precision

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    @property
    def dim(self) -> int:
        return 1 << self.n_qubits

    def reset(self) -> "DeviceState":
        _lib.check(_lib.load().qhip_state_reset(self._h))
        return self

    def upload(self, amps: np.ndarray) -> "DeviceState":
        a = np.ascontiguousarray(amps, dtype=np.complex128).ravel()
        if a.size != self.dim:
            raise ValueError(f"initial_state length {a.size} != 2^n_qubits = {self.dim}")
        _lib.check(_lib.load().qhip_state_upload_c128(self._h, a.ctypes.data, a.size))
        return self

    def download(self, out: Optional[np.ndarray] = None) -> np.ndarray:
        """The amplitudes as complex128, into `out` when given (a C-contiguous
        complex128 array of 2^n entries, written in place)."""
        if out is None:
            out = np.empty(self.dim, dtype=np.complex128)
        elif out.dtype != np.complex128 or out.size != self.dim or not out.flags.c_contiguous:
            raise ValueError("DeviceState.download: out must be C-contiguous complex128 with 2^n entries")
        _lib.check(_lib.load().qhip_state_download_c128(self._h, out.ctypes.data, out.size))
        return out

    def norm2(self) -> float:
        v = C.c_double()
        _lib.check(_lib.load().qhip_state_norm2(self._h, C.byref(v)))
        return v.value

    def copy_from(self, other: "DeviceState") -> "DeviceState":
        _lib.check(_lib.load().qhip_state_copy(self._h, other._h))
        return self

    def device_ptr(self) -> int:
        p = C.c_void_p()
        _lib.check(_lib.load().qhip_state_device_ptr(self._h, C.byref(p)))
        return p.value or 0

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib._lib is not None:
            _lib._lib.qhip_state_destroy(h)
            self._h = None


def _check_initial_state(psi, dim: int, where: str) -> np.ndarray:
    """validate_initial_state (qarp_simulator.cpp:731-746)."""
    a = np.ascontiguousarray(psi, dtype=np.complex128).ravel()
    if a.size != dim:
        raise ValueError(f"QarpSimulator::{where}: initial_state length {a.size} != 2^n_qubits = {dim}")
    norm = math.sqrt(float(np.vdot(a, a).real))
    if abs(norm - 1.0) > 1e-10:
        raise ValueError(
            f"QarpSimulator::{where}: initial_state is not normalised: |psi| = {norm:.6f} differs from 1 "
            "by more than 1e-10.  Normalise the input yourself — no silent renormalisation.")
    return a


def _random_seed() -> int:
    return secrets.randbits(32)


class QarpSimulator:
    """GPU statevector simulator with the reference ``QarpSimulator`` API."""

    MAX_FUSION_QUBITS = 8
    DEFAULT_FUSION_QUBITS = 3
    DEFAULT_FUSION_MIN_QUBITS = 12

    def __init__(self, noise_model=None, *, precision: str = "complex128", tile_qubits: int = 0,
                 low_qubits: int = 0, max_ops_per_pass: int = 0, use_mfma: int = -1, relabel_qubits: int = -1,
                 stream_tiles: int = -1):
        if precision not in ("complex128", "complex64"):
            raise ValueError(f"precision must be complex128 or complex64, got {precision!r}")
        _lib.load()
        self.noise_model = noise_model   # QarpSimulator(NoiseModel), qarp_simulator.cpp:72-77
        self.precision = precision
        self._fusion_max = _env_fusion_default()
        self._fusion_min = self.DEFAULT_FUSION_MIN_QUBITS
        self.tile_qubits = int(tile_qubits)
        self.low_qubits = int(low_qubits)
        self.max_ops_per_pass = int(max_ops_per_pass)
        self.use_mfma = int(use_mfma)   # dense k = 3..5 on the matrix cores: -1 auto, 0 never
        self.relabel_qubits = int(relabel_qubits)   # passes move qubits into the low tile positions: -1 auto, 0 never
        # Persistent prefetching tile workgroups: 1 for passes with matrix-core
        # ops, -1 only for the widest tile (2^13 complex128 / 2^14 complex64), 0 never.
        self.stream_tiles = int(stream_tiles)
        self.last_stats: Optional[dict] = None
        self._pool: dict = {}

    # ── knobs ──
    @property
    def fusion_max_qubits(self) -> int:
        return self._fusion_max

    @fusion_max_qubits.setter
    def fusion_max_qubits(self, k: int) -> None:
        k = int(k)
        if k < 0 or k > self.MAX_FUSION_QUBITS:
            raise ValueError(f"QarpSimulator::set_fusion_max_qubits: {k} exceeds the maximum block width "
                             f"{self.MAX_FUSION_QUBITS}")
        self._fusion_max = k

    @property
    def fusion_min_qubits(self) -> int:
        return self._fusion_min

    @fusion_min_qubits.setter
    def fusion_min_qubits(self, n: int) -> None:
        self._fusion_min = int(n)

    def _config(self):
        return config_struct(self._fusion_max, self._fusion_min, self.tile_qubits, self.low_qubits,
                             self.max_ops_per_pass, self.use_mfma, self.relabel_qubits, self.stream_tiles)

    # ── device-resident primitives ──
    def new_state(self, n_qubits: int, initial_state=None, where: str = "statevector") -> DeviceState:
        st = DeviceState(n_qubits, self.precision)
        if initial_state is not None:
            st.upload(_check_initial_state(initial_state, st.dim, where))
        return st

    # States of at most _POOL_BYTES that a call uses and drops (energies,
    # gradients) are kept for the next call of the same width instead of a
    # device allocation and free per call; at most two per width.  A state is
    # taken out of the pool while a call uses it, so concurrent calls never
    # share one.
    _POOL_BYTES = 1 << 30

    def _borrow_state(self, n_qubits: int, initial_state, where: str) -> DeviceState:
        pool = self._pool.get((int(n_qubits), self.precision))
        st = None
        if pool:
            try:
                st = pool.pop()
            except IndexError:
                st = None
        if st is None:
            return self.new_state(n_qubits, initial_state, where)
        if initial_state is not None:
            st.upload(_check_initial_state(initial_state, st.dim, where))
        else:
            st.reset()
        return st

    def _return_state(self, st: DeviceState) -> None:
        amp = 16 if self.precision == "complex128" else 8
        if st.dim * amp > self._POOL_BYTES:
            return
        pool = self._pool.setdefault((st.n_qubits, self.precision), [])
        if len(pool) < 2:
            pool.append(st)

    def apply(self, commands: Sequence[Command], state: DeviceState, *, terminal: bool = False,
              packed: Optional[PackedStream] = None) -> dict:
        """Run the command stream on a device state in place (the gate loop of
        QarpSimulator::statevector); returns the schedule statistics."""
        ps = packed if packed is not None else pack_stream(as_commands(commands))
        cfg = self._config()
        stats = QhipApplyStats()
        fn = _lib.load().qhip_apply_terminal if terminal else _lib.load().qhip_apply
        _lib.check(fn(C.byref(cfg), ps.ref, state.handle, C.byref(stats)))
        self.last_stats = {"fused_ops": stats.fused_ops, "passes": stats.passes,
                           "diagonal_ops": stats.diagonal_ops, "mfma_ops": stats.mfma_ops,
                           "stream_passes": stats.stream_passes, "flops": stats.flops}
        return self.last_stats

    def transition_state(self, bra: DeviceState, ket: DeviceState, observable: Observable) -> complex:
        po = pack_observable(observable)
        out = (C.c_double * 2)()
        _lib.check(_lib.load().qhip_transition(bra.handle, ket.handle, po.ref, out))
        return complex(out[0], out[1])

    def expectation_state(self, state: DeviceState, observable: Observable) -> float:
        return self.transition_state(state, state, observable).real

    def sample_state(self, state: DeviceState, n_shots: int, seed: Optional[int] = None) -> np.ndarray:
        out = np.zeros(max(int(n_shots), 0), dtype=np.uint64)
        s = _random_seed() if seed is None else int(seed) & 0xFFFFFFFF
        _lib.check(_lib.load().qhip_sample(state.handle, int(n_shots), s, out.ctypes.data if out.size else None))
        return out

    def sample_uniforms(self, state: DeviceState, uniforms: np.ndarray) -> np.ndarray:
        u = np.ascontiguousarray(uniforms, dtype=np.float64)
        out = np.zeros(u.size, dtype=np.uint64)
        _lib.check(_lib.load().qhip_sample_uniforms(state.handle, u.ctypes.data if u.size else None, u.size,
                                                    out.ctypes.data if out.size else None))
        return out

    def probabilities_state(self, state: DeviceState) -> np.ndarray:
        out = np.zeros(state.dim, dtype=np.float64)
        _lib.check(_lib.load().qhip_probabilities(state.handle, out.ctypes.data, out.size))
        return out

    def exact_state(self, state: DeviceState, threshold: float = 1e-12):
        """(keys, probs): every index with |amp|^2 > threshold, ascending, compacted on the device (the
        Born distribution of _exact_result, src/qarp/engines/_engine.py:79-116)."""
        lib = _lib.load()
        n = C.c_int64()
        _lib.check(lib.qhip_exact(state.handle, float(threshold), None, None, 0, C.byref(n)))
        keys = np.zeros(n.value, dtype=np.int64)
        probs = np.zeros(n.value, dtype=np.float64)
        if n.value:
            _lib.check(lib.qhip_exact(state.handle, float(threshold), keys.ctypes.data, probs.ctypes.data,
                                      n.value, C.byref(n)))
        return keys, probs

    # ── noise (the trajectory path only) ──
    def noise_active(self) -> bool:
        """noise_active (qarp_simulator.h:266-268): enabled and any channel."""
        return self.noise_model is not None and self.noise_model.active()

    def _refuse_noise(self, where: str, what: str) -> None:
        if self.noise_active():
            raise RuntimeError(f"QarpSimulator::{where}: noise model is active — {what}")

    # ── reference surface ──
    def statevector(self, commands: Sequence[Command], n_qubits: int, *, initial_state=None) -> np.ndarray:
        commands = as_commands(commands)
        self._refuse_noise("statevector", "the noisy ensemble is not representable as a single statevector. "
                                          "Either disable the noise model (noise_model.enabled=False) or use "
                                          "run() for trajectory simulation.")
        st = self.new_state(n_qubits, initial_state, "statevector")
        self.apply(commands, st)
        return st.download()

    def expectation(self, commands: Sequence[Command], n_qubits: int, observable: Observable, *,
                    initial_state=None) -> float:
        commands = as_commands(commands)
        self._refuse_noise("statevector", "the noisy ensemble is not representable as a single statevector.")
        st = self.new_state(n_qubits, initial_state, "expectation")
        self.apply(commands, st)
        return self.expectation_state(st, observable)

    def transition(self, bra, ket, n_qubits: int, observable: Observable) -> complex:
        n = int(n_qubits)
        if n < 0 or n > 62:
            raise ValueError(f"pauli_transition: n_qubits must be in [0, 62], got {n}")
        dim = 1 << n
        b = np.ascontiguousarray(bra, dtype=np.complex128).ravel()
        k = np.ascontiguousarray(ket, dtype=np.complex128).ravel()
        if b.size != dim or k.size != dim:
            raise ValueError(f"QarpSimulator::transition: bra/ket lengths {b.size}/{k.size} do not match "
                             f"2^{n} = {dim}")
        sb = DeviceState(n, self.precision).upload(b)
        sk = DeviceState(n, self.precision).upload(k)
        return self.transition_state(sb, sk, observable)

    def batch_expectation(self, commands: Sequence[Command], n_qubits: int, observable: Observable,
                          param_sets: Sequence[Mapping[str, float]]) -> list:
        commands = as_commands(commands)
        st = DeviceState(n_qubits, self.precision)
        po = pack_observable(observable)
        out = []
        for params in param_sets:
            st.reset()
            self.apply(commands, st, packed=pack_stream(commands, values=params))
            v = (C.c_double * 2)()
            _lib.check(_lib.load().qhip_transition(st.handle, st.handle, po.ref, v))
            out.append(v[0])
        return out

    def run(self, commands: Sequence[Command], n_qubits: int, n_shots: int, seed: Optional[int] = None, *,
            initial_state=None) -> SamplingResult:
        """Terminal-measurement fast path of QarpSimulator::run
        (qarp_simulator.cpp:1017-1036) and sample_terminal (:751-783); circuits
        with mid-circuit measurement, Reset, conditions or branches take the
        shot-batched trajectory path (trajectory.py, :1038-1211)."""
        commands = as_commands(commands)
        if self.noise_active() or not measurements_terminal(commands):
            from .trajectory import run_trajectories
            return run_trajectories(self, commands, n_qubits, n_shots, seed, initial_state,
                                    self.noise_model if self.noise_active() else None)
        st = self.new_state(n_qubits, initial_state, "run")
        self.apply(commands, st, terminal=True)
        return self._terminal_result(st, commands, n_qubits, n_shots, seed)

    def _terminal_result(self, st: DeviceState, commands: Sequence[Command], n_qubits: int, n_shots: int,
                         seed: Optional[int]) -> SamplingResult:
        outcomes = self.sample_state(st, n_shots, seed)
        res = SamplingResult(n_qubits=int(n_qubits), n_shots=int(n_shots))
        vals, cnt = np.unique(outcomes, return_counts=True)
        res.counts = {int(v): int(c) for v, c in zip(vals, cnt)}
        recorded = [(c.qubits[0], c.cbits[0]) for c in commands
                    if c.gate == GateType.Measure and c.cbits]
        res.n_cbits = cbit_register_width(commands)
        if recorded:
            hist = []
            for o in outcomes.tolist():
                reg = [False] * res.n_cbits
                for q, cb in recorded:
                    reg[cb] = bool((o >> q) & 1)
                hist.append(reg)
            res.cbit_history = hist
        return res

    # ── structured programs (QarpSimulator::execute_program / program_statevector /
    #    program_run, qarp_simulator.cpp:789-895) ──
    def _check_program_width(self, program: Program, n_qubits: int, where: str) -> None:
        w = program.min_register_width
        if int(n_qubits) < 0 or w > int(n_qubits):
            raise ValueError(f"QarpSimulator::{where}: a kernel acts on qubit {w - 1} of a "
                             f"{int(n_qubits)}-qubit register")

    def execute_program(self, program: Program, state: DeviceState, terminal_last: bool = False) -> None:
        """Run the program's kernels in order on a device state (the last gates
        kernel on the terminal-measurement path when ``terminal_last``)."""
        lib = _lib.load()
        kernels = program.kernels()
        for i, k in enumerate(kernels):
            if k[0] == "gates":
                self.apply(k[1], state, terminal=terminal_last and i + 1 == len(kernels))
            elif k[0] == "permutation":
                qs = np.ascontiguousarray(k[1], dtype=np.uint32)
                table = np.ascontiguousarray(k[2], dtype=np.uint64)
                _lib.check(lib.qhip_state_permute(state.handle, qs.ctypes.data, qs.size, table.ctypes.data))
            elif k[0] == "dense":
                self.apply([Command(GateType.Custom, qubits=k[1], unitary=k[2])], state)
            else:
                cs = np.ascontiguousarray(k[1], dtype=np.uint32)
                es = np.ascontiguousarray(k[2], dtype=np.uint64)
                ts = np.ascontiguousarray(k[3], dtype=np.uint32)
                m = np.ascontiguousarray(k[4], dtype=np.complex128)
                cfg = self._config()
                _lib.check(lib.qhip_state_controlled_powers(C.byref(cfg), state.handle, cs.ctypes.data,
                                                            es.ctypes.data, cs.size, ts.ctypes.data, ts.size,
                                                            m.ctypes.data))

    def program_statevector(self, program: Program, n_qubits: int, *, initial_state=None) -> np.ndarray:
        """statevector() for a structured program (qarp_simulator.cpp:822-849)."""
        self._check_program_width(program, n_qubits, "program_statevector")
        self._refuse_noise("program_statevector", "the noisy ensemble is not representable as a single "
                                                  "statevector.")
        for k in program.kernels():
            if k[0] == "gates" and any(needs_trajectory(c) for c in k[1]):
                raise RuntimeError("QarpSimulator::program_statevector: program contains mid-circuit "
                                   "measurement, Reset, or classical condition — statevector is not "
                                   "deterministic.")
        st = self.new_state(n_qubits, initial_state, "program_statevector")
        self.execute_program(program, st)
        return st.download()

    def program_run(self, program: Program, n_qubits: int, n_shots: int, seed: Optional[int] = None, *,
                    initial_state=None) -> SamplingResult:
        """run()'s terminal sample-once path for a structured program
        (qarp_simulator.cpp:851-888)."""
        self._check_program_width(program, n_qubits, "program_run")
        self._refuse_noise("program_run", "structured programs run the noiseless sample-once path only.")
        kernels = program.kernels()
        last: Sequence[Command] = []
        for i, k in enumerate(kernels):
            if k[0] != "gates":
                continue
            if i + 1 == len(kernels):
                last = k[1]
            elif any(needs_trajectory(c) for c in k[1]):
                raise ValueError("QarpSimulator::program_run: a measurement or classical condition "
                                 "outside the program's last gates kernel")
        if last and not measurements_terminal(last):
            raise ValueError("QarpSimulator::program_run: the last gates kernel's measurements are "
                             "not terminal")
        st = self.new_state(n_qubits, initial_state, "program_run")
        self.execute_program(program, st, terminal_last=True)
        return self._terminal_result(st, last, n_qubits, n_shots, seed)

    def make_transition_column_provider(self, U: Sequence[Command], n: int, observable,
                                        initial_batch_capacity: int = 1, sector_target_labels=None):
        """Columns U† H U |s> for batches of basis states
        (CudaqSimulator.make_transition_column_provider, bindings_montecarlo.cpp:375-381)."""
        from .columns import TransitionColumnProvider
        return TransitionColumnProvider(self, U, n, observable, initial_batch_capacity, sector_target_labels)

    def batch_run(self, commands: Sequence[Command], n_qubits: int, n_shots: int,
                  param_sets: Sequence[Mapping[str, float]], seed: Optional[int] = None) -> list:
        commands = as_commands(commands)
        out = []
        for i, params in enumerate(param_sets):
            s = None if seed is None else (int(seed) + i) & 0xFFFFFFFF
            concrete = [c.substitute(params) for c in commands]
            out.append(self.run(concrete, n_qubits, n_shots, s))
        return out

    def run_gradient(self, commands: Sequence[Command], n_qubits: int, observable: Observable,
                     params: Mapping[str, float], param_order: Sequence[str], *, initial_state=None) -> list:
        """Adjoint gradient of <H> (run_gradient, qarp_simulator.cpp:1607-1624)."""
        grad, _ = self.run_gradient_energy(commands, n_qubits, observable, params, param_order,
                                           initial_state=initial_state)
        return grad.tolist()

    def run_gradient_energy(self, commands, n_qubits, observable, params, param_order, *,
                            initial_state=None, state: Optional[DeviceState] = None):
        """(gradient array, energy) in one adjoint sweep."""
        commands = as_commands(commands)
        index = {name: i for i, name in enumerate(param_order)}
        ps = pack_stream(commands, values=params, param_index=index)
        po = pack_observable(observable)
        st = state if state is not None else self._borrow_state(n_qubits, initial_state, "run_gradient")
        try:
            grad = np.zeros(len(param_order))
            energy = C.c_double()
            cfg = self._config()
            _lib.check(_lib.load().qhip_run_gradient(C.byref(cfg), ps.ref, len(param_order), st.handle, po.ref,
                                                     grad.ctypes.data if grad.size else None, C.byref(energy)))
        finally:
            if state is None:
                self._return_state(st)
        return grad, energy.value

    def run_gradient_phi(self, commands: Sequence[Command], n_qubits: int, phi, params: Mapping[str, float],
                         param_order: Sequence[str], *, initial_state=None) -> list:
        commands = as_commands(commands)
        index = {name: i for i, name in enumerate(param_order)}
        ps = pack_stream(commands, values=params, param_index=index)
        st = self.new_state(n_qubits, initial_state, "run_gradient")
        ph = np.ascontiguousarray(np.asarray(phi, dtype=np.complex128).ravel())
        if ph.size != st.dim:
            raise RuntimeError("run_gradient: |φ⟩ size mismatch (expected 2^n_qubits)")
        sp = DeviceState(n_qubits, self.precision).upload(ph)
        grad = np.zeros(len(param_order))
        cfg = self._config()
        _lib.check(_lib.load().qhip_run_gradient_phi(C.byref(cfg), ps.ref, len(param_order), st.handle,
                                                     sp.handle, grad.ctypes.data if grad.size else None))
        return grad.tolist()

    def batch_energy_grad(self, instances: Sequence, gradients: bool = True):
        """Energies (and adjoint gradients) of many independent instances in
        batched launches (qhip_batch_energy_grad).  ``instances`` holds tuples
        (commands, n_qubits, observable, params, param_order).  Returns
        (energies array, list of gradient arrays or None)."""
        b = _PackedBatch(instances)
        energies = np.zeros(b.m)
        grads = np.zeros(max(b.total, 1)) if gradients else None
        cfg = self._config()
        _lib.check(_lib.load().qhip_batch_energy_grad(
            C.byref(cfg), b.m, b.nq.ctypes.data, b.streams, b.observables, b.npar.ctypes.data,
            QHIP_C128 if self.precision == "complex128" else QHIP_C64, energies.ctypes.data,
            grads.ctypes.data if gradients else None, b.offs.ctypes.data))
        return energies, b.split(grads) if gradients else None

    def fused_stream(self, commands: Sequence[Command], n_qubits: int) -> list:
        """The fusion pass output: list of (gate id, qubits, params, matrix-or-None)."""
        ps = pack_stream(as_commands(commands))
        cfg = self._config()
        lib = _lib.load()
        h = C.c_void_p()
        _lib.check(lib.qhip_fuse(C.byref(cfg), ps.ref, int(n_qubits), C.byref(h)))
        try:
            n = C.c_int64()
            _lib.check(lib.qhip_fused_count(h, C.byref(n)))
            out = []
            qs = np.zeros(64, np.uint32)
            pr = np.zeros(4)
            m = np.zeros(256 * 256, np.complex128)
            for i in range(n.value):
                g, k = C.c_int32(), C.c_int32()
                _lib.check(lib.qhip_fused_get(h, i, C.byref(g), C.byref(k), qs.ctypes.data, pr.ctypes.data,
                                              m.ctypes.data, m.size))
                mat = None
                if g.value == int(GateType.Custom) and k.value > 0:
                    d = 1 << k.value
                    mat = m[: d * d].reshape(d, d).copy()
                out.append((g.value, tuple(int(q) for q in qs[: k.value]), tuple(pr), mat))
            return out
        finally:
            lib.qhip_fused_destroy(h)


class _PackedBatch:
    """Instances (commands, n_qubits, observable, params, param_order) as the
    C structs of qhip_batch_* see them (buffers kept alive here)."""

    def __init__(self, instances: Sequence):
        from ._abi import QhipObservable
        self.instances = [(as_commands(c), n, o, p, order) for c, n, o, p, order in instances]
        self.m = len(self.instances)
        self.nq = np.zeros(self.m, dtype=np.int32)
        self.npar = np.zeros(self.m, dtype=np.int32)
        self.offs = np.zeros(self.m, dtype=np.int64)
        total = 0
        self.packed_o = []
        for i, (cmds, n, obs, params, order) in enumerate(self.instances):
            self.packed_o.append(pack_observable(obs))
            self.nq[i], self.npar[i], self.offs[i] = int(n), len(order), total
            total += len(order)
        self.total = total
        self.observables = (QhipObservable * max(self.m, 1))(*[p.struct for p in self.packed_o])
        self.bind([inst[3] for inst in self.instances])

    def bind(self, values: Sequence[Mapping[str, float]]) -> None:
        from ._abi import QhipStream
        self.packed_s = []
        for (cmds, n, obs, _, order), vals in zip(self.instances, values):
            index = {name: k for k, name in enumerate(order)}
            self.packed_s.append(pack_stream(cmds, values=vals, param_index=index))
        self.streams = (QhipStream * max(self.m, 1))(*[p.struct for p in self.packed_s])

    def split(self, grads: np.ndarray) -> list:
        return [grads[self.offs[i]:self.offs[i] + self.npar[i]].copy() for i in range(self.m)]


class BatchSession:
    """Independent instances kept on the device for parameter sweeps
    (qhip_batch_*): ``run`` evaluates energies (and adjoint gradients) of
    every instance, replaying the sweep as one captured hipGraph when
    ``use_graph``; ``update(values)`` rebinds one parameter dict per
    instance (True when the session had to be rebuilt)."""

    def __init__(self, sim: QarpSimulator, instances: Sequence, gradients: bool = True):
        self.sim = sim
        self.gradients = bool(gradients)
        self._b = _PackedBatch(instances)
        self._h = None
        h = C.c_void_p()
        cfg = sim._config()
        _lib.check(_lib.load().qhip_batch_create(
            C.byref(cfg), self._b.m, self._b.nq.ctypes.data, self._b.streams, self._b.observables,
            self._b.npar.ctypes.data, QHIP_C128 if sim.precision == "complex128" else QHIP_C64,
            int(self.gradients), C.byref(h)))
        self._h = h

    def run(self, use_graph: bool = True):
        energies = np.zeros(self._b.m)
        grads = np.zeros(max(self._b.total, 1)) if self.gradients else None
        _lib.check(_lib.load().qhip_batch_run(self._h, energies.ctypes.data,
                                              grads.ctypes.data if self.gradients else None,
                                              self._b.offs.ctypes.data, int(bool(use_graph))))
        return energies, self._b.split(grads) if self.gradients else None

    def stats(self) -> dict:
        """Algorithmic HBM bytes of one sweep from the session's schedules
        (qhip_batch_stats): forward passes, H|psi> and the energies, reverse
        two-state passes; and the tile-pass launch counts."""
        from ._abi import QhipBatchInfo
        info = QhipBatchInfo()
        _lib.check(_lib.load().qhip_batch_stats(self._h, C.byref(info)))
        return {name: getattr(info, name) for name, _ in QhipBatchInfo._fields_}

    def update(self, values: Sequence[Mapping[str, float]]) -> bool:
        if len(values) != self._b.m:
            raise ValueError(f"update: expected {self._b.m} parameter dicts, got {len(values)}")
        self._b.bind(values)
        rebuilt = C.c_int32(0)
        _lib.check(_lib.load().qhip_batch_update(self._h, self._b.streams, C.byref(rebuilt)))
        return bool(rebuilt.value)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib._lib is not None:
            _lib._lib.qhip_batch_destroy(h)
            self._h = None


class CompiledCircuit:
    """A command stream fused and scheduled once, op tables resident on the
    device (qhip_plan_*); ``apply`` launches only the passes, ``apply_graph``
    replays them as one captured hipGraph, and ``update(values)`` rebinds the
    symbolic parameters of the same circuit (rewriting only the coefficient
    table while the schedule keeps its structure)."""

    def __init__(self, sim: "QarpSimulator", commands: Sequence[Command], n_qubits: int, *,
                 fusion_max_qubits: Optional[int] = None, tile_qubits: Optional[int] = None,
                 low_qubits: Optional[int] = None, use_mfma: Optional[int] = None,
                 relabel_qubits: Optional[int] = None, values: Optional[Mapping[str, float]] = None,
                 stream_tiles: Optional[int] = None):
        self.n_qubits = int(n_qubits)
        self.precision = sim.precision
        cfg = config_struct(sim.fusion_max_qubits if fusion_max_qubits is None else int(fusion_max_qubits),
                            sim.fusion_min_qubits,
                            sim.tile_qubits if tile_qubits is None else int(tile_qubits),
                            sim.low_qubits if low_qubits is None else int(low_qubits),
                            sim.max_ops_per_pass, sim.use_mfma if use_mfma is None else int(use_mfma),
                            sim.relabel_qubits if relabel_qubits is None else int(relabel_qubits),
                            sim.stream_tiles if stream_tiles is None else int(stream_tiles))
        self.config = {"fusion_max_qubits": cfg.fusion_max_qubits, "tile_qubits": cfg.tile_qubits,
                       "low_qubits": cfg.low_qubits, "use_mfma": cfg.use_mfma,
                       "relabel_qubits": cfg.relabel_qubits, "stream_tiles": cfg.stream_tiles}
        self._cfg = cfg
        self._commands = as_commands(list(commands))
        ps = pack_stream(self._commands, values=values)
        self._h = None
        h = C.c_void_p()
        _lib.check(_lib.load().qhip_plan_create(C.byref(cfg), ps.ref, self.n_qubits,
                                                QHIP_C128 if self.precision == "complex128" else QHIP_C64,
                                                C.byref(h)))
        self._h = h
        self.stats: dict = {}     # filled by the first apply

    def apply(self, state: DeviceState) -> dict:
        stats = QhipApplyStats()
        _lib.check(_lib.load().qhip_plan_apply(self._h, state.handle, C.byref(stats)))
        self.stats = {"fused_ops": stats.fused_ops, "passes": stats.passes,
                      "diagonal_ops": stats.diagonal_ops, "mfma_ops": stats.mfma_ops,
                      "stream_passes": stats.stream_passes, "flops": stats.flops}
        return self.stats

    def apply_graph(self, state: DeviceState) -> dict:
        stats = QhipApplyStats()
        _lib.check(_lib.load().qhip_plan_apply_graph(self._h, state.handle, C.byref(stats)))
        self.stats = {"fused_ops": stats.fused_ops, "passes": stats.passes,
                      "diagonal_ops": stats.diagonal_ops, "mfma_ops": stats.mfma_ops,
                      "stream_passes": stats.stream_passes, "flops": stats.flops}
        return self.stats

    def update(self, values: Mapping[str, float]) -> bool:
        """Rebind parameter values; True when the device tables had to be
        rebuilt (the schedule changed structure), False when only the
        coefficients were rewritten."""
        ps = pack_stream(self._commands, values=values)
        rebuilt = C.c_int32(0)
        _lib.check(_lib.load().qhip_plan_update(self._h, C.byref(self._cfg), ps.ref, C.byref(rebuilt)))
        return bool(rebuilt.value)

    def apply_timed(self, state: DeviceState) -> np.ndarray:
        """Apply once; per-pass kernel durations (ms) from HIP events recorded
        on the library's stream around every pass launch."""
        if not self.stats:
            self.apply(state)
        n = int(self.stats["passes"])
        out = np.zeros(max(n, 1))
        _lib.check(_lib.load().qhip_plan_apply_timed(self._h, state.handle, out.ctypes.data, n))
        return out[:n]

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib._lib is not None:
            _lib._lib.qhip_plan_destroy(h)
            self._h = None


def inspect_schedule(commands: Sequence[Command], n_qubits: int, *, fusion_max_qubits: int = 3,
                     fusion_min_qubits: int = 12, tile_qubits: int = 0, low_qubits: int = 0,
                     precision: str = "complex128", relabel_qubits: int = -1) -> list:
    """The device schedule of ``commands`` as [(pass, Command)] in execution
    order, computed on the host only (qhip_schedule_inspect): dense and
    diagonal blocks as Custom, mask phases as MCZ, scalar phases as GPhase,
    on the qubit positions the kernels address.  A pass that relabels qubits
    ends with the SWAP gates that realise its store permutation, so applying
    the list in order reproduces the circuit's state."""
    lib = _lib.load()
    cfg = config_struct(fusion_max_qubits, fusion_min_qubits, tile_qubits, low_qubits, 0, -1, relabel_qubits)
    ps = pack_stream(as_commands(commands))
    h = C.c_void_p()
    _lib.check(lib.qhip_schedule_inspect(C.byref(cfg), ps.ref, int(n_qubits),
                                         QHIP_C128 if precision == "complex128" else QHIP_C64, C.byref(h)))
    try:
        n = C.c_int64()
        _lib.check(lib.qhip_fused_count(h, C.byref(n)))
        out = []
        qs = np.zeros(64, np.uint32)
        pr = np.zeros(4)
        m = np.zeros(256 * 256, np.complex128)
        for i in range(n.value):
            g, k, p = C.c_int32(), C.c_int32(), C.c_int32()
            _lib.check(lib.qhip_fused_get(h, i, C.byref(g), C.byref(k), qs.ctypes.data, pr.ctypes.data,
                                          m.ctypes.data, m.size))
            _lib.check(lib.qhip_fused_pass(h, i, C.byref(p)))
            q = [int(x) for x in qs[: k.value]]
            if g.value == int(GateType.GPhase):
                cmd = Command(GateType.GPhase, params=[math.atan2(pr[1], pr[0])])
            elif g.value == int(GateType.MCZ):
                if abs(complex(pr[0], pr[1]) + 1) > 1e-15:
                    raise AssertionError("mask phase other than -1")
                cmd = Command(GateType.MCZ, qubits=q)
            else:
                d = 1 << k.value
                cmd = Command(GateType.Custom, qubits=q, unitary=m[: d * d].reshape(d, d).copy())
            out.append((p.value, cmd))
        return out
    finally:
        lib.qhip_fused_destroy(h)
