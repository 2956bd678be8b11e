#!/usr/bin/env python3
"""Benchmark of the MI355X statevector backend on one GPU.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Headline workload: 30-qubit random brickwork circuit in complex128 (16 GiB
state): 30 layers, each a seeded random pairing of the 30 qubits carrying a
Haar-random 2-qubit unitary (450 gates), gate fusion up to width 5 (width and
tile shape picked by the autotuner during warmup).  A step is one full
circuit apply on the state resident in HBM.  value = HBM passes x (read +
write of the state) / time per step, in GB/s.

Also on the JSON line:
  roofline      dominant kernel (the tile pass): algorithmic bytes per launch
                over its mean launch time from HIP events on the library's
                stream, against the 8 TB/s HBM3E peak;
  expectation   <H> of a 2000-string random Pauli sum on the 30-qubit state;
  c1            20-qubit hardware-efficient ansatz: state, <H> (TFIM), 4000 shots;
  adjoint       24-qubit, 384-parameter VQE energy gradient: adjoint sweep vs
                a parameter-shift sweep (768 shifted energies) on the GPU;
  batched_vqe   512 independent instances, energies + adjoint gradients in
                batched launches, complex128 and complex64 (error per instance);
  qaoa          34-qubit complex64 QAOA MaxCut: <C>, 1e5 shots, the adjoint
                gradient at 33 qubits and complex64 vs complex128 at 32;
  cpu_baseline  the reference C++ simulator (compiled from its sources into
                oracle/_ref) on bounded samples of the same workloads, on
                this host's cores, medians of three runs.

    python bench.py --cpu-protocol   # the reference build alone, every leg at
                                     # its stated size (no GPU used)
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "fused-apply HBM GB/s (% peak) at 30q; adjoint VQE grads/s vs param-shift"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cores() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, int(os.environ.get("OMP_NUM_THREADS", n))))


def fused_apply(args, sim):
    from openqarp_amd import workloads as W
    from openqarp_amd.autotune import tune
    from openqarp_amd.simulator import CompiledCircuit

    n = args.qubits
    cmds = W.brickwork_haar(n, args.layers, seed=2)
    state = sim.new_state(n)
    log(f"brickwork n={n} layers={args.layers}: {len(cmds)} gates; autotuning")
    widths = (args.fusion,) if args.fusion else (2, 3, 4, 5)
    tuned = tune(sim, cmds, state, widths=widths, tiles=(args.tile,) if args.tile else None,
                 lows=(args.low,) if args.low else None, try_stream=not args.no_stream)
    best = tuned["best"]
    for t in tuned["trials"]:
        log(f"  trial k={t['fusion_max_qubits']} T={t['tile_qubits']} low={t['low_qubits']} "
            f"stream={t['stream_tiles']}: "
            f"{t['ms']:.1f} ms, {t['passes']} passes, {t['fused_ops']} fused ops")
    log(f"autotune best {best}")
    cc = CompiledCircuit(sim, cmds, n, fusion_max_qubits=best["fusion_max_qubits"],
                         tile_qubits=best["tile_qubits"], low_qubits=best["low_qubits"],
                         stream_tiles=best["stream_tiles"])
    state.reset()
    for _ in range(args.warmup):
        cc.apply(state)
    # The library runs on its own HIP stream; qhip_synchronize drains it.
    from openqarp_amd import _lib
    _lib.check(_lib.load().qhip_synchronize())
    pass_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        pass_ms.append(cc.apply_timed(state))
    _lib.check(_lib.load().qhip_synchronize())
    wall = (time.perf_counter() - t0) / args.steps
    passes = int(cc.stats["passes"])
    stream_passes = int(cc.stats["stream_passes"])
    kernel = ("tile_stream_kernel" if stream_passes == passes else "tile_pass_kernel" if stream_passes == 0
              else f"tile_stream_kernel ({stream_passes} passes) + tile_pass_kernel")
    bytes_per_pass = 2 * 16 * (1 << n)
    value = passes * bytes_per_pass / wall / 1e9
    per_pass = np.concatenate(pass_ms)
    kernel_ms = float(per_pass.mean())
    achieved = bytes_per_pass / (kernel_ms * 1e-3) / 1e9
    norm = state.norm2()
    log(f"apply: {wall * 1e3:.2f} ms/step, {passes} passes ({stream_passes} streaming), {value:.1f} GB/s, "
        f"kernel mean {kernel_ms:.3f} ms, norm2 {norm:.15f}")
    # Streaming floor: one tile pass carrying a single 1-qubit gate (same tile
    # shape), and the driver's device-to-device copy of the state, as the
    # read + write references for the per-pass bandwidth.
    if args.no_floor:
        return {
            "state": state, "n": n, "wall_ms": wall * 1e3, "value": value, "passes": passes,
            "stream_passes": stream_passes, "kernel": kernel,
            "fused_ops": int(cc.stats["fused_ops"]), "gates": len(cmds), "best": best,
            "trials": tuned["trials"], "kernel_ms": kernel_ms, "achieved": achieved,
            "gate_equiv_gbs": len(cmds) * bytes_per_pass / wall / 1e9, "norm2": norm,
            "flops": float(cc.stats["flops"]), "pass_floor": None, "device_copy": None,
        }
    from openqarp_amd.ir import Command, GateType
    one = CompiledCircuit(sim, [Command(GateType.H, 0)], n, fusion_max_qubits=best["fusion_max_qubits"],
                          tile_qubits=best["tile_qubits"], low_qubits=best["low_qubits"],
                          stream_tiles=best["stream_tiles"])
    one.apply(state)
    floor_ms = float(np.concatenate([one.apply_timed(state) for _ in range(5)]).mean())
    other = sim.new_state(n)
    lib = _lib.load()
    _lib.check(lib.qhip_state_copy(other._h, state._h))
    t0 = time.perf_counter()
    for _ in range(5):
        _lib.check(lib.qhip_state_copy(other._h, state._h))
    copy_ms = (time.perf_counter() - t0) / 5 * 1e3
    del other
    floor_gbs = bytes_per_pass / (floor_ms * 1e-3) / 1e9
    copy_gbs = bytes_per_pass / (copy_ms * 1e-3) / 1e9
    log(f"one-gate tile pass {floor_ms:.3f} ms = {floor_gbs:.1f} GB/s; device copy {copy_ms:.3f} ms = {copy_gbs:.1f} GB/s")
    return {
        "state": state, "n": n, "wall_ms": wall * 1e3, "value": value, "passes": passes,
        "stream_passes": stream_passes, "kernel": kernel,
        "fused_ops": int(cc.stats["fused_ops"]), "gates": len(cmds), "best": best,
        "trials": tuned["trials"], "kernel_ms": kernel_ms, "achieved": achieved,
        "gate_equiv_gbs": len(cmds) * bytes_per_pass / wall / 1e9, "norm2": norm,
        "flops": float(cc.stats["flops"]),
        "pass_floor": {"ms": floor_ms, "GBps": floor_gbs, "frac": floor_gbs / HBM_PEAK_GBS,
                       "what": "one tile pass carrying a single H gate, same tile shape"},
        "device_copy": {"ms": copy_ms, "GBps": copy_gbs, "what": "qhip_state_copy (hipMemcpy device to device)"},
    }


def expectation(sim, state, n):
    from openqarp_amd import workloads as W
    obs = W.random_pauli_sum(n, 2000, seed=3)
    sim.expectation_state(state, obs[:5])   # warm
    t0 = time.perf_counter()
    e = sim.expectation_state(state, obs)
    dt = time.perf_counter() - t0
    groups = len({tuple(sorted(q for q, l in p if l in "XY")) for p, _ in obs})
    log(f"<H> 2000 strings ({groups} x-groups): {dt * 1e3:.1f} ms, value {e:.12f}")
    return {"terms": len(obs), "x_groups": groups, "ms": dt * 1e3, "value": e}


def adjoint(sim, n=24, layers=8, points=None):
    from openqarp_amd import workloads as W
    from openqarp_amd.engine import HipEngine
    from openqarp_amd.primitives import StateVector
    w = W.c3_gradient(n, layers, seed=3, n_terms=1000)
    names = w.symbols
    sim.run_gradient_energy(w.commands, n, w.observable, w.values, names)   # warm
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        grad, energy = sim.run_gradient_energy(w.commands, n, w.observable, w.values, names)
    t_adj = (time.perf_counter() - t0) / reps
    log(f"adjoint {n}q {len(names)} params: {t_adj * 1e3:.1f} ms, E={energy:.12f}")
    # Parameter shift through the engine: every occurrence shifted on its own
    # (two-term rule for Ry/Rz, src/qarp/engines/_gradients.py:250-301), all
    # 768 shifted energies through the batched forward + <H> launches.
    eng = HipEngine(simulator=sim)
    prim = StateVector(w.commands, n, w.observable)
    sub = names if points is None else names[:max(1, points // 2)]
    t0 = time.perf_counter()
    ps = eng.run_gradient(prim, w.values, sub, method="parameter-shift")
    t_sub = time.perf_counter() - t0
    t_ps = t_sub * len(names) / len(sub)
    err = float(np.abs(ps - grad[:len(sub)]).max())
    log(f"parameter shift: {2 * len(sub)}/{2 * len(names)} points in {t_sub:.2f} s -> {t_ps:.2f} s per gradient, "
        f"max |adj - ps| = {err:.3e}")
    return {"n_qubits": n, "params": len(names), "terms": len(w.observable), "adjoint_ms": t_adj * 1e3,
            "adjoint_grads_per_s": 1.0 / t_adj, "param_shift_s": t_ps, "param_shift_points_timed": 2 * len(sub),
            "param_shift_points": 2 * len(names), "param_shift_grads_per_s": 1.0 / t_ps,
            "speedup_vs_param_shift": t_ps / t_adj, "max_abs_diff_vs_param_shift": err, "energy": energy}


def batched(n_instances=512):
    """512 independent HEA/TFIM instances (12..22 qubits, depth 4..16): energies
    and adjoint gradients in batched launches, complex128 then complex64 with
    FP64 reductions, and the complex64 error per instance."""
    from openqarp_amd import workloads as W
    from openqarp_amd.simulator import QarpSimulator
    ws = W.batched_vqe(n_instances, seed=4)
    inst = [(w.commands, w.n_qubits, w.observable, w.values, w.symbols) for w in ws]
    out = {"instances": n_instances, "params": int(sum(len(w.symbols) for w in ws)),
           "state_gib_c128": round(sum(16 << w.n_qubits for w in ws) / 2**30, 2)}
    res = {}
    for prec in ("complex128", "complex64"):
        sim = QarpSimulator(precision=prec)
        sim.batch_energy_grad(inst[:8])   # warm
        t0 = time.perf_counter()
        e, g = sim.batch_energy_grad(inst)
        dt = time.perf_counter() - t0
        res[prec] = (e, g)
        out[f"{prec}_s"] = dt
        log(f"batched {n_instances} instances {prec}: {dt:.3f} s")
    # Optimizer-style sweeps: the same instances kept in a session, the sweep
    # replayed as one hipGraph after each coefficient update.
    from openqarp_amd.simulator import BatchSession
    sess = BatchSession(QarpSimulator(), inst)
    sess.run(use_graph=True)   # capture
    rng = np.random.default_rng(1)
    vals = [{s: float(rng.uniform(0, 2 * math.pi)) for s in w.symbols} for w in ws]
    t0 = time.perf_counter()
    sess.update(vals)
    t_upd = time.perf_counter() - t0
    t0 = time.perf_counter()
    sess.run(use_graph=True)
    t_graph = time.perf_counter() - t0
    t0 = time.perf_counter()
    sess.run(use_graph=False)
    t_direct = time.perf_counter() - t0
    out.update({"session_update_s": t_upd, "session_sweep_graph_s": t_graph, "session_sweep_direct_s": t_direct})
    log(f"session sweep: update {t_upd:.3f} s, graph replay {t_graph:.3f} s, direct launches {t_direct:.3f} s")
    # Roofline of the sweep: algorithmic bytes from the session's schedules
    # (forward passes 2·B·2^n, H|psi> gathers, reverse two-state passes
    # 4·B·2^n per instance) over the graph-replayed sweep time.
    info = sess.stats()
    total = info["fwd_bytes"] + info["obs_bytes"] + info["rev_bytes"]
    gbs = total / t_graph / 1e9
    out["sweep_roofline"] = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": round(gbs / HBM_PEAK_GBS, 4), "bytes": total, **info}
    log(f"session sweep: {total / 1e9:.1f} GB algorithmic ({info['fwd_launches']} forward + "
        f"{info['rev_launches']} reverse launches) in {t_graph:.3f} s = {gbs:.0f} GB/s")
    del sess
    e128, g128 = res["complex128"]
    e64, g64 = res["complex64"]
    de = np.abs(e64 - e128)
    dg = np.array([np.abs(a - b).max() if a.size else 0.0 for a, b in zip(g64, g128)])
    out.update({"c64_energy_abs_err_max": float(de.max()), "c64_energy_abs_err_median": float(np.median(de)),
                "c64_grad_abs_err_max": float(dg.max()), "c64_grad_abs_err_median": float(np.median(dg))})
    log(f"complex64 vs complex128: energy err max {de.max():.2e} median {np.median(de):.2e}; "
        f"grad err max {dg.max():.2e}")
    return out


def qaoa(n=34, p=4, seed=5, shots=100_000, n_adj=33, n_cmp=32, fusion=1, low=3):
    """QAOA MaxCut on a seeded random 3-regular graph: <C> and 1e5 shots on a
    34-qubit complex64 state (128 GiB), the adjoint gradient at 33 qubits in
    complex64, and complex64 vs complex128 at 32 qubits.  The apply fuses
    single-qubit runs only, so each pass's share of a cost layer (RZZ gates,
    consecutive) is one merged phase op, with 3 contiguous low qubits: 12
    passes, 1.41 s, against 1.89 s at fusion width 2 (17 passes of dense
    2-qubit ops), 3.9 s at width 4 and 5.3 s at width 3."""
    import gc
    from openqarp_amd import workloads as W
    from openqarp_amd.simulator import DeviceState, QarpSimulator
    from openqarp_amd import _lib
    from openqarp_amd.simulator import CompiledCircuit
    sim = QarpSimulator(precision="complex64", low_qubits=low)
    sim.fusion_max_qubits = fusion
    w = W.qaoa_maxcut(n, p, seed)
    st = DeviceState(n, "complex64")
    # Fused and scheduled once; one untimed apply, then the timed one from |0>.
    t0 = time.perf_counter()
    cc = CompiledCircuit(sim, w.commands, n)
    t_plan = time.perf_counter() - t0
    cc.apply(st)
    st.reset()
    _lib.check(_lib.load().qhip_synchronize())
    t0 = time.perf_counter()
    stats = cc.apply(st)
    _lib.check(_lib.load().qhip_synchronize())
    t_apply = time.perf_counter() - t0
    del cc
    t0 = time.perf_counter()
    c = sim.expectation_state(st, w.observable)
    t_c = time.perf_counter() - t0
    t0 = time.perf_counter()
    outcomes = sim.sample_state(st, shots, seed)
    t_s = time.perf_counter() - t0
    edges = [(t[0][0], t[1][0]) for t, _ in w.observable]
    zz = np.zeros(outcomes.size)
    for a, b in edges:
        zz += 1.0 - 2.0 * (((outcomes >> a) ^ (outcomes >> b)) & 1).astype(np.float64)
    c_shots = float(zz.mean())
    del st
    gc.collect()
    log(f"qaoa {n}q c64: plan {t_plan:.2f} s, apply {t_apply:.2f} s ({stats['passes']} passes), <C> {c:.6f} in {t_c:.3f} s, "
        f"{shots} shots in {t_s:.2f} s (shot mean {c_shots:.4f})")
    wa = W.qaoa_maxcut(n_adj, p, seed, symbolic=True)
    t0 = time.perf_counter()
    ga, ea = sim.run_gradient_energy(wa.commands, n_adj, wa.observable, wa.values, wa.symbols)
    t_adj = time.perf_counter() - t0
    gc.collect()
    log(f"qaoa {n_adj}q c64 adjoint: {t_adj:.2f} s, E {ea:.6f}")
    wc = W.qaoa_maxcut(n_cmp, p, seed, symbolic=True)
    g64, e64 = sim.run_gradient_energy(wc.commands, n_cmp, wc.observable, wc.values, wc.symbols)
    gc.collect()
    g128, e128 = QarpSimulator().run_gradient_energy(wc.commands, n_cmp, wc.observable, wc.values, wc.symbols)
    gc.collect()
    log(f"qaoa {n_cmp}q: c64 vs c128 energy err {abs(e64 - e128):.2e}, grad err {np.abs(g64 - g128).max():.2e}")
    return {"n_qubits": n, "p": p, "edges": len(edges), "state_gib": (8 << n) / 2**30,
            "apply_s": t_apply, "plan_s": t_plan, "passes": stats["passes"],
            "apply_GBps": stats["passes"] * 2 * 8 * (1 << n) / t_apply / 1e9, "expectation": c, "expectation_s": t_c,
            "shots": shots, "sample_s": t_s, "shot_mean_C": c_shots,
            "adjoint_n_qubits": n_adj, "adjoint_s": t_adj, "adjoint_energy": ea,
            "compare_n_qubits": n_cmp, "c64_energy_abs_err": abs(e64 - e128),
            "c64_grad_abs_err": float(np.abs(g64 - g128).max())}


def structured(sim, state, n):
    """Structured-program kernels on the apply leg's state: basis
    permutations of the 20 lowest qubits (gathered reads inside 2^20-amplitude
    blocks) and of the 10 highest (whole 16-MiB runs move), and QPE-style
    controlled powers (8 controls with exponents 2^j, 2 targets; one control,
    10 targets through the controlled dense kernel).  Bytes per
    permutation: one out-of-place pass, 2 x 16 B x 2^n (the state takes the
    permuted buffer)."""
    from openqarp_amd import _lib
    from openqarp_amd.program import Program
    from openqarp_amd.workloads import haar_unitary
    rng = np.random.default_rng(9)
    out = {}
    for name, qs in (("permutation_low20", list(range(20))), ("permutation_high10", list(range(n - 10, n)))):
        p = Program()
        p.add_permutation(qs, rng.permutation(1 << len(qs)))
        sim.execute_program(p, state)
        t0 = time.perf_counter()
        for _ in range(3):
            sim.execute_program(p, state)
        ms = (time.perf_counter() - t0) / 3 * 1e3
        gbs = 2 * 16 * (1 << n) / (ms * 1e-3) / 1e9
        out[name] = {"ms": round(ms, 3), "GBps": round(gbs, 1)}
        log(f"{name}: {ms:.2f} ms = {gbs:.0f} GB/s (one out-of-place pass)")
    p = Program()
    controls = list(range(n - 1, n - 9, -1))
    p.add_controlled_powers(controls, [1 << j for j in range(8)], [0, 1], haar_unitary(4, rng))
    sim.execute_program(p, state)
    _lib.check(_lib.load().qhip_synchronize())
    t0 = time.perf_counter()
    sim.execute_program(p, state)
    _lib.check(_lib.load().qhip_synchronize())
    ms = (time.perf_counter() - t0) * 1e3
    out["controlled_powers_8x2"] = {"ms": round(ms, 3)}
    log(f"controlled powers (8 controls, 2 targets): {ms:.2f} ms")
    # Ten targets: wider than one tile-pass op, so the controlled dense
    # kernel runs; 8 flops per complex multiply-add, 2^(n-1) x 2^10 of them.
    p = Program()
    p.add_controlled_powers([n - 1], [1], list(range(10)), haar_unitary(1 << 10, rng))
    sim.execute_program(p, state)
    _lib.check(_lib.load().qhip_synchronize())
    t0 = time.perf_counter()
    sim.execute_program(p, state)
    _lib.check(_lib.load().qhip_synchronize())
    ms = (time.perf_counter() - t0) * 1e3
    tf = 8.0 * 2.0 ** (n - 1) * 1024 / (ms * 1e-3) / 1e12
    out["controlled_power_1x10"] = {"ms": round(ms, 3), "tflops_f64": round(tf, 2)}
    log(f"controlled power (1 control, 10 targets): {ms:.2f} ms = {tf:.1f} TFLOP/s f64")
    # Phase-estimation shape: four controls with exponents 1, 2, 4, 8 on the
    # same ten targets: three squarings of the 1024 x 1024 matrix on the
    # device, then four controlled GEMMs.
    p = Program()
    p.add_controlled_powers([n - 4, n - 3, n - 2, n - 1], [1, 2, 4, 8], list(range(10)), haar_unitary(1 << 10, rng))
    sim.execute_program(p, state)
    _lib.check(_lib.load().qhip_synchronize())
    t0 = time.perf_counter()
    sim.execute_program(p, state)
    _lib.check(_lib.load().qhip_synchronize())
    ms = (time.perf_counter() - t0) * 1e3
    tf = (4 * 8.0 * 2.0 ** (n - 1) * 1024 + 3 * 8.0 * 1024.0 ** 3) / (ms * 1e-3) / 1e12
    out["controlled_powers_4x10"] = {"ms": round(ms, 3), "tflops_f64": round(tf, 2)}
    log(f"controlled powers (4 controls, exponents 1..8, 10 targets): {ms:.2f} ms = {tf:.1f} TFLOP/s f64")
    return out


def columns(n=20, capacity=32, n_sources=128, n_terms=400):
    """Transition columns U† H U |s> (TransitionColumnProvider): a 20-qubit
    brickwork U, a 400-term local Pauli sum, 128 sources in batches of 32
    (one 25-qubit register per batch)."""
    from openqarp_amd import workloads as W
    from openqarp_amd.simulator import QarpSimulator
    sim = QarpSimulator()
    U = W.brickwork_haar(n, 10, 6)
    obs = [(t, float(np.real(c))) for t, c in W.random_local_pauli_sum(n, n_terms, 7)]
    prov = sim.make_transition_column_provider(U, n, obs, capacity)
    src = np.arange(0, n_sources * 4099, 4099, dtype=np.uint64) % np.uint64(1 << n)
    prov.columns(src[:capacity])
    t0 = time.perf_counter()
    prov.compact_columns(src, 1e-12)
    s = time.perf_counter() - t0
    log(f"transition columns: {n_sources} columns of 2^{n} in {s:.3f} s ({n_sources / s:.0f} columns/s)")
    d = prov.diagnostics
    return {"n_qubits": n, "sources": n_sources, "batch": capacity, "terms": n_terms, "s": round(s, 4),
            "columns_per_s": round(n_sources / s, 1),
            "u_s": round(d["u_wall_seconds"], 4), "h_s": round(d["hamiltonian_wall_seconds"], 4),
            "u_dagger_s": round(d["u_dagger_wall_seconds"], 4), "transfer_s": round(d["transfer_wall_seconds"], 4)}


def c1(reps=5):
    """The 20-qubit hardware-efficient ansatz (10 layers of RY, RZ and a CX
    ladder) in complex128 from |0>: the state (downloaded to the host, as
    the reference returns it), <H> for the open-chain TFIM and 4000 shots,
    median of `reps` runs after one warm-up."""
    from openqarp_amd import workloads as W
    from openqarp_amd.simulator import QarpSimulator
    w = W.c1_hea(20, 10, seed=1)
    sim = QarpSimulator()

    def run():
        st = sim.new_state(20)
        sim.apply(w.commands, st)
        e = sim.expectation_state(st, w.observable)
        shots = sim.sample_state(st, 4000, 1234)
        return st.download(), e, shots
    run()
    dt, ts = _median_time(run, reps)
    sv, e, _ = run()
    log(f"C1 20q HEA: state + <H> + 4000 shots in {dt * 1e3:.2f} ms (median of {reps}), <H> = {e:.12f}")
    return {"n_qubits": 20, "layers": 10, "s": dt, "runs_s": ts, "energy": e, "norm2": float(np.vdot(sv, sv).real),
            "what": "statevector (downloaded) + <H> + 4000 shots, complex128"}


def trajectories(n=16, shots=2048):
    """Trajectory path: a 16-qubit brickwork, mid-circuit measurements with
    a conditional correction and a reset, more brickwork, 2048 shots in
    batches of 256 (one 24-qubit state per batch).  The reference build runs
    the same circuit and shots in 5.05 s on an 8-core host."""
    from openqarp_amd import workloads as W
    from openqarp_amd.ir import Command, GateType
    from openqarp_amd.simulator import QarpSimulator
    sim = QarpSimulator()
    cmds = W.brickwork_haar(n, 4, 11) + [
        Command(GateType.Measure, 0, cbits=[0]),
        Command(GateType.X, 1, condition_bits=[0], condition_values=[True]),
        Command(GateType.Reset, 2),
    ] + W.brickwork_haar(n, 4, 12) + [Command(GateType.Measure, 3, cbits=[1])]
    sim.run(cmds, n, 64, seed=1)
    t0 = time.perf_counter()
    res = sim.run(cmds, n, shots, seed=2)
    s = time.perf_counter() - t0
    log(f"trajectories: {shots} shots of a {n}-qubit circuit with mid-circuit measurement in {s:.3f} s")
    return {"n_qubits": n, "shots": shots, "s": round(s, 4), "shots_per_s": round(shots / s, 1),
            "distinct_outcomes": len(res.counts)}


def _median_time(fn, reps):
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), ts


def cpu_baseline(full=False, reps=3):
    """The reference simulator compiled from its sources (oracle/_ref) on the
    host's cores, timed as medians of `reps` runs.  Default sizes are bounded
    samples of the benchmark workloads (a few tens of seconds in total); with
    full=True every leg runs at its stated size (C2 apply at 30 qubits, C3 at
    24), which takes most of an hour on 8 cores.

    Legs: apply (brickwork, GB/s in the headline's unit), c1 (20-qubit HEA:
    statevector, <H>, 4000 shots), c2_expectation (<H> of the 2000-string sum
    on a brickwork state), c3_energy (forward energy of the CZ-brickwork
    ansatz with 1000 local strings; one parameter-shift point is two of
    these), c3_adjoint (the reference's adjoint gradient)."""
    cores = host_cores()
    os.environ["QARP_NUM_THREADS"] = str(cores)
    os.environ["OMP_NUM_THREADS"] = str(cores)
    sys.path.insert(0, str(ROOT / "tests"))
    import _oracle as O
    from openqarp_amd import workloads as W
    from openqarp_amd._abi import pack_observable, pack_stream
    if O.reference() is None:
        # Without the reference build: the C restatement, single-threaded.
        n, fusion = 20, 3
        ps = pack_stream(W.brickwork_haar(n, 30, seed=2))
        nf = len(O.fuse(ps, n, fusion, 12))
        dt, ts = _median_time(lambda: O.statevector(ps, n, fusion, 12), reps)
        gbs = nf * 2 * 16 * (1 << n) / dt / 1e9
        log(f"cpu baseline (port) n={n}: median {dt:.2f} s, {nf} fused passes, {gbs:.2f} GB/s")
        return {"value": gbs, "unit": "GB/s", "cores": 1, "kind": "port",
                "sample": f"brickwork n={n} layers=30 fusion k={fusion}, {nf} fused passes, median of {reps} "
                          f"runs {dt:.2f} s; GB/s = passes x 2 x 16 B x 2^n / time"}
    legs = {}
    # Headline family: brickwork at fusion width 3 (the reference's default).
    n = 30 if full else 24
    ps = pack_stream(W.brickwork_haar(n, 30, seed=2))
    nf = len(O.ref_fuse(ps, n, 3, 12))
    dt, ts = _median_time(lambda: O.statevector(ps, n, 3, 12, lib="ref"), reps)
    gbs = nf * 2 * 16 * (1 << n) / dt / 1e9
    legs["apply"] = {"n_qubits": n, "passes": nf, "s": dt, "runs_s": ts, "GBps": gbs}
    log(f"cpu baseline (reference) brickwork n={n}: median {dt:.2f} s, {nf} fused passes, {gbs:.2f} GB/s")
    # C1 in full.
    w = W.c1_hea(20, 10, seed=1)
    ps1, po1 = pack_stream(w.commands), pack_observable(w.observable)

    def c1():
        sv = O.statevector(ps1, 20, 3, 12, lib="ref")
        O.transition(sv, sv, 20, po1, lib="ref")
        O.sample(sv, 20, 4000, 1234, lib="ref")
    dt, ts = _median_time(c1, reps)
    legs["c1"] = {"n_qubits": 20, "s": dt, "runs_s": ts, "what": "statevector + <H> + 4000 shots"}
    log(f"cpu baseline (reference) C1 20q: median {dt:.3f} s")
    # C2's <H> (2000 strings) on a brickwork state.
    n2 = 30 if full else 22
    w2 = W.c2_brickwork(n2, 30, seed=2)
    sv2 = O.statevector(pack_stream(w2.commands), n2, 3, 12, lib="ref")
    po2 = pack_observable(w2.observable)
    # At 30 qubits one run takes about half an hour on 8 cores: timed once.
    dt, ts = _median_time(lambda: O.transition(sv2, sv2, n2, po2, lib="ref"), 1 if full else reps)
    legs["c2_expectation"] = {"n_qubits": n2, "terms": len(w2.observable), "s": dt, "runs_s": ts}
    log(f"cpu baseline (reference) C2 <H> {n2}q: median {dt:.2f} s")
    del sv2
    # C3: forward energy (one parameter-shift point is two of them) and adjoint.
    n3 = 24 if full else 22
    g = W.c3_gradient(n3, 8, seed=3, n_terms=1000)
    idx = {s: i for i, s in enumerate(g.symbols)}
    vals = np.array([g.values[s] for s in g.symbols])
    ps3, po3 = pack_stream(g.commands, g.values, idx), pack_observable(g.observable)

    def energy():
        sv = O.statevector(ps3, n3, 3, 12, lib="ref")
        return O.transition(sv, sv, n3, po3, lib="ref")
    dt, ts = _median_time(energy, reps)
    legs["c3_energy"] = {"n_qubits": n3, "params": len(vals), "terms": 1000, "s": dt, "runs_s": ts,
                         "param_shift_gradient_s_est": 2 * len(vals) * dt}
    log(f"cpu baseline (reference) C3 energy {n3}q: median {dt:.2f} s (a parameter-shift gradient: "
        f"{2 * len(vals)} of them)")
    na = 24 if full else 20
    ga = W.c3_gradient(na, 8, seed=3, n_terms=1000)
    ida = {s: i for i, s in enumerate(ga.symbols)}
    va = np.array([ga.values[s] for s in ga.symbols])
    psa, poa = pack_stream(ga.commands, ga.values, ida), pack_observable(ga.observable)
    ra = reps if full else 1
    dt, ts = _median_time(lambda: O.run_gradient(psa, na, poa, len(va), values=va, lib="ref"), ra)
    legs["c3_adjoint"] = {"n_qubits": na, "params": len(va), "terms": 1000, "s": dt, "runs_s": ts,
                          "grads_per_s": 1.0 / dt}
    log(f"cpu baseline (reference) C3 adjoint {na}q, {len(va)} params: {dt:.2f} s")
    return {"value": legs["apply"]["GBps"], "unit": "GB/s", "cores": cores, "kind": "reference",
            "sample": f"brickwork n={n} layers=30 fusion k=3 (the reference's default width), {nf} fused passes, "
                      f"median of {reps} runs {legs['apply']['s']:.2f} s; GB/s = passes x 2 x 16 B x 2^n / time",
            "legs": legs}


def traffic_from_profiles():
    p = ROOT / "profiles" / "tile_pass_traffic.json"
    if p.exists():
        try:
            return json.loads(p.read_text()).get("bytes_per_launch")
        except (ValueError, OSError):
            return None
    return None


LEGS = ("apply", "expectation", "structured", "c1", "adjoint", "batch", "qaoa", "columns", "trajectories", "cpu")


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--layers", type=int, default=30)
    ap.add_argument("--fusion", type=int, default=0, help="fix the fusion width (0: autotune 2..5)")
    ap.add_argument("--tile", type=int, default=0, help="fix the tile qubits (0: autotune)")
    ap.add_argument("--low", type=int, default=0, help="fix the contiguous low qubits (0: autotune)")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-expectation", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--ps-points", type=int, default=0, help="time only this many parameter-shift points")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-qaoa", action="store_true")
    ap.add_argument("--no-structured", action="store_true",
                    help="skip the structured-program kernels, transition columns and trajectories")
    ap.add_argument("--no-floor", action="store_true", help="skip the one-gate pass and device-copy references")
    ap.add_argument("--no-stream", action="store_true",
                    help="run every pass one tile per workgroup (no persistent prefetching tile kernel)")
    ap.add_argument("--legs", default="all",
                    help="comma-separated subset of " + ",".join(LEGS) + " (default: all)")
    ap.add_argument("--cpu-protocol", action="store_true",
                    help="time only the reference build, every leg at its stated size (no GPU; most of an hour)")
    args = ap.parse_args(argv)
    if args.cpu_protocol:
        print(json.dumps({"cpu_protocol": cpu_baseline(full=True)}), flush=True)
        return 0
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported (the backend is single-GPU; "
              "the multi-GPU layout is a design note, not a build target)", file=sys.stderr)
        return 2

    legs = set(LEGS if args.legs == "all" else args.legs.split(","))
    unknown = legs - set(LEGS)
    if unknown:
        print(f"unknown legs {sorted(unknown)}; choose from {','.join(LEGS)}", file=sys.stderr)
        return 2
    skip = {"expectation": args.no_expectation, "structured": args.no_structured, "adjoint": args.no_adjoint,
            "batch": args.no_batch, "qaoa": args.no_qaoa, "columns": args.no_structured,
            "trajectories": args.no_structured, "cpu": args.no_cpu}
    legs = {leg for leg in legs if not skip.get(leg, False)}

    from openqarp_amd.simulator import QarpSimulator
    sim = QarpSimulator(stream_tiles=0 if args.no_stream else -1)
    need_state = bool(legs & {"apply", "expectation", "structured"})
    fa = fused_apply(args, sim) if need_state else None
    ex = expectation(sim, fa["state"], fa["n"]) if "expectation" in legs else None
    st = structured(sim, fa["state"], fa["n"]) if "structured" in legs else None
    if fa is not None:
        fa["state"] = None
    adj = adjoint(sim, points=args.ps_points or None) if "adjoint" in legs else None
    bat = batched() if "batch" in legs else None
    qa = qaoa() if "qaoa" in legs else None
    cols = columns() if "columns" in legs else None
    traj = trajectories() if "trajectories" in legs else None
    hea = c1() if "c1" in legs else None
    cpu = cpu_baseline() if "cpu" in legs else None
    if fa is None:
        print(json.dumps({"metric": METRIC, "legs": sorted(legs), "c1": hea, "expectation": ex,
                          "structured_program": st, "transition_columns": cols, "trajectories": traj, "adjoint": adj,
                          "batched_vqe": bat, "qaoa": qa, "cpu_baseline": cpu}), flush=True)
        return 0
    line = {
        "metric": METRIC, "value": round(fa["value"], 2), "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(fa["wall_ms"], 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "c128 (f64)", "data": "synthetic",
        "config": {"workload": f"brickwork{fa['n']}-haar2q-c128", "n_qubits": fa["n"], "layers": args.layers,
                   "gates": fa["gates"], "fusion_max_qubits": fa["best"]["fusion_max_qubits"],
                   "tile_qubits": fa["best"]["tile_qubits"], "low_qubits": fa["best"]["low_qubits"],
                   "stream_tiles": fa["best"]["stream_tiles"],
                   "fused_ops": fa["fused_ops"], "passes": fa["passes"], "stream_passes": fa["stream_passes"],
                   "parallelism": "single-gpu"},
        "pct_of_peak": round(100.0 * fa["value"] / HBM_PEAK_GBS, 2),
        "gate_equivalent_GBps": round(fa["gate_equiv_gbs"], 1),
        "fused_gates_per_pass": round(fa["fused_ops"] / max(fa["passes"], 1), 2),
        "roofline": {"bound": "hbm", "achieved": round(fa["achieved"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(fa["achieved"] / HBM_PEAK_GBS, 4), "traffic": traffic_from_profiles(),
                     "kernel": fa["kernel"], "mean_launch_ms": round(fa["kernel_ms"], 4),
                     "bytes_per_launch": 2 * 16 * (1 << fa["n"])},
        "pass_floor": fa["pass_floor"], "device_copy": fa["device_copy"],
        "autotune": fa["trials"],
        "c1": hea, "expectation": ex, "structured_program": st, "transition_columns": cols, "trajectories": traj,
        "adjoint": adj,
        "batched_vqe": bat, "qaoa": qa,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
