#!/usr/bin/env python3
"""FCIQMC step throughput on one MI355X.

Workload: HubbardReal1D, M = N = 20 (periodic), u = 6, t = 1, IsDynamicSemistochastic
(FP64 amplitudes, ThresholdCompression(1), rel = 1, abs = Inf), DoubleLogUpdate with
target_walkers = 1e8, dtau = 1e-3, seed 17, starting from near_uniform with population 10
and shift = H_aa = 0 (Rimu's defaults).  The population is first grown to the target with
the same engine (equilibration, not timed), then `--warmup` steps, then `--steps` timed
steps.  A "step" is one apply_operator! + compress! + norm/len of the whole vector.

value = spawn attempts per second over the timed steps (Rimu's `spawn_attempts` stat).
The roofline object refers to the kernel with the largest device time; its algorithmic
bytes are the records it must read and write once.  cpu_baseline times the threaded C++
restatement of Rimu's PDVec step (oracle/) on a bounded sample of the same state.

Two more objects on the same line:
* "batched": 256 HubbardReal1D runs of a U-sweep advanced together for 10,000 steps
  (solves/s), against the plain engine running the same runs one at a time; each run's
  energy estimates are compared with its own Lanczos ground-state energy.
* "adjoint": dE/du and dE/dt of the K = 2000-step deterministic projector energy for
  HubbardReal1D M = N = 12 at u = 2, 4, 6, 8 in one product (grads/s), against central
  differences on the plain engine.
* "momentum": the momentum-space case, HubbardMom1D M = N = 16, u = 4, integer walkers,
  2e7 walkers, dtau = 5e-4: spawn attempts/s and per-kernel times at steady state.
* "initiator": the initiator approximation.  (a) Rimu's benchmark case "(10, 20) Mom space
  with projected energy and initiator" (benchmark/benchmarks.jl:58-66): HubbardMom1D with
  10 bosons in mode 10 of 20, u = 1, PDVec(addr => 1.0; IsDynamicSemistochastic,
  initiator = true), ProjectedEnergy, DoubleLogUpdate(target_walkers = 40_000),
  dtau = 1e-4, 8000 steps through lomc, timed whole; steady-state step times of the engine
  and of the oracle's initiator step (one core) from its final vector.  (b) The headline
  state (1e8 walkers) stepped with Initiator(1): per-kernel times next to the plain step.
* "checkpoint": save_state / load_state of the headline state as a Rimu Arrow state file
  (download + zstd IPC write; read + upload), in GB/s of state.
* "overlaps": the AllOverlaps reductions of the headline state v against the state one step
  later w: dot(v, w), dot(v, G2RealCorrelator(1), w), dot(v, H, w).
* "wide": two-chunk addresses (99 bits).  (a) Rimu's benchmark case "(50, 50) Real space"
  (benchmark/benchmarks.jl:66-73): near_uniform(BoseFS{50,50}), HubbardReal1D u = 6,
  PDVec(addr => 1.0; IsDynamicSemistochastic), DoubleLogUpdate(target_walkers = 50_000),
  dtau = 1e-4, 1000 steps through lomc, timed whole; steady-state step times of the engine
  and of the oracle (one core) from its final vector.  (b) The same model grown to 1e8
  walkers with the headline's settings: spawn attempts/s and per-kernel times.
* "realspace": HubbardRealSpace.  (a) 2D Bose-Hubbard, 64 bosons on PeriodicBoundaries(8, 8),
  u = 6 (127-bit addresses), grown to 1e8 walkers: spawn attempts/s, per-kernel times and
  the oracle's step on a sample.  (b) The reference's "2D Hubbard" system
  (benchmark/benchmarks.jl:8-14, FermiFS2C on PeriodicBoundaries(4, 4)) projected to its
  ground state with IsDeterministic (2000 steps), energy and time.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
from __graft_entry__ import load_package  # noqa: E402

HBM_PEAK = 8.0e12  # B/s, MI355X

# HBM bytes per launch measured on this workload (target 1e8 walkers, the equilibrated state:
# about 7.07e7 spawn records onto 3.42e7 entries) with rocprofv3 FETCH_SIZE / WRITE_SIZE, scaled
# by the same counters on a plain 1 GiB copy (fetch x2.0, write x1.0), expressed as a ratio to
# the kernel's algorithmic bytes.  radix_count is the one histogram pass over the keys,
# radix_scatter the look-back scatter, compact the placement of the bucket survivors.
MEASURED_TRAFFIC_RATIO = {"bucket_merge": 1.019, "radix_scatter": 1.200, "spawn": 1.162, "plan": 1.016,
                          "compact": 1.027, "radix_count": 0.830, "merge": 1.042}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_alg_bytes(name, S, L, Lnew, vsize=8, ksize=8):
    """Algorithmic bytes of one launch: every record the kernel must touch, read or
    written once.  Records are (key, 8-byte value); keys are 8 bytes (one-chunk addresses)
    or 16 bytes (two chunks)."""
    rec = ksize + vsize
    if name == "radix_scatter":
        return 2 * S * rec
    if name == "radix_count":
        return S * ksize
    if name in ("merge", "bucket_merge"):
        return S * rec + L * (ksize + vsize) + Lnew * rec
    if name == "compact":
        return 2 * Lnew * rec
    if name == "spawn":
        return L * (rec + 8) + S * rec
    if name == "plan":
        return L * rec + L * (vsize + 8)
    return None


def run_cpu_baseline(rimu_keys, rimu_vals, args, step_no, shift):
    """Threaded PDVec-style restatement (oracle/liboracle.so) on a sample of the state."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    threads = args.cpu_threads or min(16, os.cpu_count() or 1)
    n = len(rimu_keys)
    stride = max(1, n // args.cpu_sample_entries)
    keys = np.ascontiguousarray(rimu_keys[::stride])
    vals = np.ascontiguousarray(rimu_vals[::stride])
    model = O.real1d(20, 20, 6.0, 1.0)
    sty = O.style("dynamic")
    attempts = 0.0
    t0 = time.perf_counter()
    steps = 0
    k, v = keys, vals
    while steps < args.cpu_sample_steps:
        k, v, st = O.step(model, sty, k, v, seed=17, step=step_no + steps, shift=shift,
                          dtau=1e-3, threads=threads)
        attempts += st["attempts"]
        steps += 1
    dt = time.perf_counter() - t0
    return {"value": attempts / dt, "unit": "spawn_attempts/s", "cores": threads,
            "kind": "port",
            "sample": (f"{steps} steps of the threaded PDVec-step restatement on every "
                       f"{stride}-th entry of the equilibrated GPU state ({len(keys)} entries, "
                       f"{float(np.abs(vals).sum()):.3g} walkers), {dt:.1f} s"),
            "seconds": dt}


def _finite(x):
    """The result with every non-finite float replaced by None (strict JSON)."""
    if isinstance(x, dict):
        return {k: _finite(v) for k, v in x.items()}
    if isinstance(x, (list, tuple)):
        return [_finite(v) for v in x]
    if isinstance(x, (float, np.floating)):
        return float(x) if math.isfinite(float(x)) else None
    if isinstance(x, np.integer):
        return int(x)
    return x


def run_batched_leg(rimu, args):
    """The batched U-sweep: 256 HubbardReal1D runs (M ~ U{6..14}, N ~ U{ceil(M/2)..M},
    u ~ U[1,10], target walkers log-uniform in [1e3, 1e5], IsStochasticInteger, dtau 1e-3,
    seeds 0..255) for `--batched-steps` steps in one launch per step, graph-replayed.
    solves/s = runs completed per second of the whole sweep.  The plain engine's figure is
    the same sweep run one instance at a time: each instance's steady-state step time on
    the plain engine (from the batched run's final vector and shift) times the step count."""
    specs = rimu.u_sweep_specs(args.batched_runs)
    nsteps = args.batched_steps
    # autotune on a copy that has grown to steady state (snapshot/restore inside)
    tune = rimu.BatchedRuns(specs, history_steps=args.batched_tune_at)
    tune.step(args.batched_tune_at, steps_per_graph=100)
    blk, rad, pers, table = tune.autotune(trial_steps=20)
    tune.close()
    log(f"batched autotune: block {blk} radix {rad} {'persistent launch' if pers else 'five launches'} per step")
    b = rimu.BatchedRuns(specs, history_steps=nsteps)
    b.set_config(blk, rad, persistent=pers)
    t0 = time.perf_counter()
    b.step(nsteps, steps_per_graph=100)
    elapsed = time.perf_counter() - t0
    st, ln = b.status()
    hist = b.history()
    attempts = float(hist["spawn_attempts"].sum())
    log(f"batched sweep: {len(specs)} runs x {nsteps} steps in {elapsed:.2f} s")

    ed = batched_ed_check(rimu, specs, hist, st, skip=nsteps // 4, details=args.batched_details)

    plain_s = 0.0
    plain_per_run = {}
    sample_steps = args.plain_sample_steps
    for r, sp in enumerate(specs):
        d = sp.descriptor()
        k, v = b.vector(r)
        if st[r] != 0 or len(k) == 0:
            continue
        eng = rimu._native.Engine(sp.hamiltonian.descriptor(), rimu.IsStochasticInteger().descriptor(),
                                  d.seed, d.max_length + 1, int(min(4 * d.target_walkers + 4096, 2 ** 32 - 2)))
        eng.upload(k, v)
        shift = float(hist["shift"][-1, r])
        for i in range(3):
            eng.step(nsteps + 1 + i, shift, d.dtau)
        t1 = time.perf_counter()
        for i in range(sample_steps):
            eng.step(nsteps + 4 + i, shift, d.dtau)
        plain_per_run[r] = (time.perf_counter() - t1) / sample_steps * nsteps
        plain_s += plain_per_run[r]
        eng.close()
    b.close()
    # the extrapolation checked end to end on a sample of the runs: each sampled run from its
    # start address through all nsteps on the plain engine (host DoubleLogUpdate), timed whole
    sample = list(range(0, len(specs), max(1, len(specs) // max(1, args.plain_e2e_runs))))[:args.plain_e2e_runs]
    e2e_s = est_s = 0.0
    for r in sample:
        sp = specs[r]
        d = sp.descriptor()
        if st[r] != 0:
            continue
        eng = rimu._native.Engine(sp.hamiltonian.descriptor(), rimu.IsStochasticInteger().descriptor(),
                                  d.seed, d.max_length + 1, int(min(4 * d.target_walkers + 4096, 2 ** 32 - 2)))
        eng.upload(np.array([d.start_address], np.uint64), np.array([d.start_population]))
        strat = rimu.DoubleLogUpdate(target_walkers=d.target_walkers, zeta=d.zeta, xi=d.xi)
        shp = rimu.strategies.DefaultShiftParameters(d.shift, float(abs(d.start_population)), d.dtau)
        t1 = time.perf_counter()
        for i in range(1, nsteps + 1):
            stp = eng.step(i, shp.shift, d.dtau)
            strat.update(shp, stp.norm)
        e2e_s += time.perf_counter() - t1
        est_s += plain_per_run[r]
        eng.close()
    solves = len(specs) / elapsed
    plain = len(specs) / plain_s
    return {"metric": "solves/s", "value": solves, "plain_value": plain, "speedup": solves / plain,
            "runs": len(specs), "steps": nsteps, "seconds": elapsed,
            "spawn_attempts_per_s": attempts / elapsed,
            "status_counts": {rimu.batched.STATUS[c]: int((st == c).sum()) for c in np.unique(st)},
            "ed_check": ed,
            "autotune": {"block": blk, "radix_bits": rad, "persistent": pers,
                         "tuned_after_steps": args.batched_tune_at,
                         "ms_per_step": {f"{'persistent' if k[0] else 'five launches'} {k[1]}x{k[2]}":
                                         (float(v) if np.isfinite(v) else None) for k, v in table.items()}},
            "plain_estimate": (f"each run's steady-state step time on the plain engine "
                               f"({sample_steps} steps from the sweep's final vector) x {nsteps} steps, "
                               f"summed over runs: {plain_s:.1f} s"),
            "plain_end_to_end_check": {"runs": sample, "measured_s": e2e_s, "estimated_s": est_s,
                                       "measured_over_estimated": e2e_s / est_s if est_s else None,
                                       "speedup_scaled_to_measured": (solves / plain) * (e2e_s / est_s)
                                       if est_s else None}}


def batched_ed_check(rimu, specs, hist, status, skip, details=None):
    """Each run's shift and projected-energy estimates (blocking analysis, ratio_of_means
    with Monte Carlo error propagation, after `skip` steps) against its own ground-state
    energy (Lanczos on the device, runs of equal M, N grouped eight to a launch).  Not part
    of the timed sweep."""
    B = rimu.batched
    t0 = time.perf_counter()
    e0 = B.ground_state_energies(specs)
    t_ed = time.perf_counter() - t0
    rows = B.energy_check(specs, hist, status, e0, skip=skip)
    if details:
        with open(details, "w") as f:
            json.dump(rows, f)
    n = len(rows)
    return {"runs_checked": n, "ed_seconds": t_ed, "skip_steps": skip,
            "shift_within_2sigma": sum(r["shift_2sigma"] for r in rows),
            "shift_within_3sigma": sum(r["shift_3sigma"] for r in rows),
            "projected_within_2sigma": sum(r["projected_2sigma"] for r in rows),
            "projected_within_3sigma": sum(r["projected_3sigma"] for r in rows),
            "projected_delta_y_at_least_0.1": sum(abs(r["delta_y"]) >= 0.1 for r in rows),
            "binomial_2sigma_expectation": 0.9545 * n,
            "median_abs_shift_error_over_E0": float(np.median([abs(r["shift"] - r["E0"]) / abs(r["E0"])
                                                              for r in rows])) if rows else None}


def run_momentum_leg(rimu, args):
    """HubbardMom1D M = N = 16, u = 4, t = 1 from the zero-momentum condensate (16 bosons
    in mode 8), IsStochasticInteger, DoubleLogUpdate to 2e7 walkers, dtau = 5e-4, seed 17:
    grown to the target (untimed), then `--mom-steps` steady-state steps timed."""
    M = N = 16
    onr = [0] * M
    onr[7] = N
    H = rimu.HubbardMom1D(rimu.BoseFS(*onr), u=4.0, t=1.0)
    style = rimu.IsStochasticInteger()
    target, dtau = args.mom_target, 5e-4
    eng = rimu._native.Engine(H.descriptor(), style.descriptor(), 17, int(2 * target + 101),
                              int(min(4 * target + 4096, 2 ** 32 - 2)))
    start = H.starting_address()
    eng.upload(np.array([start.bits], dtype=np.uint64), np.array([10], dtype=np.int64))
    strat = rimu.DoubleLogUpdate(target_walkers=target)
    sp = rimu.strategies.DefaultShiftParameters(H.diagonal_element(start), 10.0, dtau)
    step, settle = 0, None
    t0 = time.perf_counter()
    while step < args.mom_equilibration:
        step += 1
        st = eng.step(step, sp.shift, dtau)
        strat.update(sp, st.norm)
        if settle is None and st.norm >= 0.9 * target:
            settle = step
        if settle is not None and step >= settle + args.mom_settle:
            break
    log(f"momentum: grown in {step} steps ({time.perf_counter() - t0:.1f} s), len {st.len:.3g} norm {st.norm:.3g}")
    eng.set_timing(True)
    attempts = alg_bytes = 0.0
    S_tot = L_tot = Lnew_tot = 0.0
    L_prev = len(eng)
    t1 = time.perf_counter()
    for _ in range(args.mom_steps):
        step += 1
        st = eng.step(step, sp.shift, dtau)
        strat.update(sp, st.norm)
        attempts += st.spawn_attempts
        b, records = eng.last_traffic()
        alg_bytes += b
        S_tot += records - L_prev
        L_tot += L_prev
        Lnew_tot += st.len
        L_prev = st.len
    elapsed = time.perf_counter() - t1
    times = eng.kernel_times()
    eng.close()
    K = args.mom_steps
    roofline, kern = kernel_rooflines(times, S_tot / K, L_tot / K, Lnew_tot / K, K)
    return {"workload": "HubbardMom1D M=N=16 u=4 t=1, condensate start, IsStochasticInteger, "
                        "target_walkers=2e7, dtau=5e-4, seed 17",
            "metric": "spawn_attempts/s", "value": attempts / elapsed,
            "ms_per_step": 1e3 * elapsed / K, "steps": K,
            "growth_steps": step - K, "walkers": st.norm, "len": st.len,
            "spawn_records_per_step": S_tot / K,
            "step_alg_roofline_frac": alg_bytes / elapsed / HBM_PEAK,
            "roofline": roofline, "shift": sp.shift, "kernels": kern}


def kernel_rooflines(times, S, L, Lnew, steps, traffic_ratio=None):
    """Per-kernel algorithmic GB/s from the engine's HIP-event times (S records, L entries in,
    Lnew out per step), and the roofline of the kernel with the most device time among those
    with an algorithmic byte count."""
    kernels = {}
    for name, (ms, launches) in times.items():
        kernels[name] = {"ms_total": ms, "launches": launches, "ms_avg": ms / max(launches, 1),
                         "launches_per_step": launches / steps}
        bpl = kernel_alg_bytes(name, S, L, Lnew)
        if bpl is not None:
            kernels[name]["alg_GBps"] = bpl / (kernels[name]["ms_avg"] / 1e3) / 1e9
    known = [n for n in times if kernel_alg_bytes(n, 1, 1, 1) is not None]
    if not known:
        return None, kernels
    dom = max(known, key=lambda n: times[n][0])
    ms, launches = times[dom]
    bytes_per_launch = kernel_alg_bytes(dom, S, L, Lnew)
    avg_s = ms / launches / 1e3
    achieved = bytes_per_launch / avg_s
    traffic = None
    if traffic_ratio and dom in traffic_ratio:
        traffic = traffic_ratio[dom] * bytes_per_launch
    roofline = {"bound": "hbm", "kernel": dom, "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": traffic,
                "alg_bytes_per_launch": bytes_per_launch, "avg_launch_ms": avg_s * 1e3,
                "launches_per_step": launches / steps}
    return roofline, kernels


def run_checkpoint_leg(rimu, eng, M, N, step, shift):
    """save_state / load_state of the equilibrated headline state (RimuIO.jl:92-180): device
    download, Arrow table with the BoseFS extension type, zstd IPC file; then read back,
    compare and upload into the engine."""
    import shutil
    import tempfile
    import pyarrow.ipc as ipc
    from rimu_amd import rimuio
    d = tempfile.mkdtemp(prefix="rimu_ckpt_")
    path = os.path.join(d, "state.arrow")
    try:
        t0 = time.perf_counter()
        keys, vals = eng.download()
        t1 = time.perf_counter()
        table = rimuio.state_table(M, N, keys, vals, step=step, shift=shift)
        with ipc.new_file(path, table.schema,
                          options=ipc.IpcWriteOptions(compression="zstd", use_threads=True)) as w:
            w.write_table(table)
        t2 = time.perf_counter()
        _, _, k2, v2, meta = rimuio.read_state_arrays(path)
        t3 = time.perf_counter()
        eng.upload(k2, v2)
        t4 = time.perf_counter()
        size = os.path.getsize(path)
        same = bool(np.array_equal(keys, k2) and np.array_equal(vals.view(np.uint64), v2.view(np.uint64)))
        del table
    finally:
        shutil.rmtree(d, ignore_errors=True)
    nbytes = 16.0 * len(keys)
    return {"entries": len(keys), "state_bytes": nbytes, "file_bytes": size,
            "download_s": t1 - t0, "write_s": t2 - t1, "read_s": t3 - t2, "upload_s": t4 - t3,
            "save_GBps": nbytes / (t2 - t0) / 1e9, "load_GBps": nbytes / (t4 - t2) / 1e9,
            "round_trip_bit_exact": same, "meta_step": meta.get("step")}


def run_overlap_leg(rimu, eng, H, step, shift, args):
    """AllOverlaps' reductions at the headline scale: the equilibrated state v and the state
    one step later w (a second engine), dot(v, w), dot(v, G2RealCorrelator(1), w) and
    dot(v, H, w) on the device (rimu_overlaps), timed with HIP events."""
    keys, vals = eng.download()
    target = args.target
    eng2 = rimu._native.Engine(H.descriptor(), rimu.IsDynamicSemistochastic().descriptor(), 17,
                               int(round(2 * target + 100)) + 1, int(min(4 * target + 4096, 2 ** 32 - 2)))
    eng2.upload(keys, vals)
    eng2.step(step + 1, shift, 1e-3)
    L1, L2 = len(eng), len(eng2)
    wk, _ = eng2.download()
    lookups = float(2 * np.bitwise_count(wk & ~(wk << np.uint64(1))).astype(np.int64).sum())
    del keys, vals, wk
    out = {"len_v": L1, "len_w": L2}
    cases = [("dot", []), ("g2", [rimu.G2RealCorrelator(1)]), ("hamiltonian", [H])]
    for name, ops in cases:
        descs = [rimu.observables.operator_descriptor(o, H.M, H.N) for o in ops]
        rimu._native.overlaps(eng, eng2, descs)           # warm-up
        eng2.set_timing(True)
        reps = args.overlap_reps if name != "hamiltonian" else 2
        t0 = time.perf_counter()
        for _ in range(reps):
            vals = rimu._native.overlaps(eng, eng2, descs)
        wall = (time.perf_counter() - t0) / reps
        times = eng2.kernel_times()
        eng2.set_timing(False)
        ms, launches = times.get("overlap", (float("nan"), 1))
        kernel_ms = ms / max(launches, 1)
        rec = {"values": [float(x) for x in vals], "kernel_ms": kernel_ms, "call_ms": 1e3 * wall}
        alg = 16.0 * (L1 + L2)   # every entry of both vectors read once
        if name != "hamiltonian":
            rec["alg_bytes"] = alg
            rec["alg_GBps"] = alg / (kernel_ms / 1e3) / 1e9
            rec["hbm_frac"] = rec["alg_GBps"] * 1e9 / HBM_PEAK
        else:
            rec["neighbour_lookups"] = lookups
            rec["lookups_per_s"] = lookups / (kernel_ms / 1e3)
        out[name] = rec
    eng2.close()
    return out


def run_initiator_case(rimu, args):
    """benchmark/benchmarks.jl:58-66 through lomc, then steady-state step times."""
    M, N = 20, 10
    onr = [0] * M
    onr[9] = N
    addr = rimu.BoseFS(*onr)
    H = rimu.HubbardMom1D(addr, u=1.0)
    dv = rimu.PDVec((addr, 1.0), style=rimu.IsDynamicSemistochastic(), initiator=True)
    post = rimu.ProjectedEnergy(H, dv)
    s_strat = rimu.DoubleLogUpdate(target_walkers=40_000)
    dtau, last = 1e-4, args.init_laststep
    t0 = time.perf_counter()
    df, sim = rimu.lomc(H, dv, s_strat=s_strat, post_step=post, dtau=dtau, laststep=last,
                        random_seed=17)
    run_s = time.perf_counter() - t0
    shift = float(df["shift"].to_numpy()[-1])
    half = len(df) // 2
    e_shift = float(np.mean(df["shift"].to_numpy()[half:]))
    e_proj = float(np.sum(df["hproj"].to_numpy()[half:]) / np.sum(df["vproj"].to_numpy()[half:]))
    keys, vals = sim.engine.download()
    eng = sim.engine
    eng.set_timing(True)
    n_gpu = args.init_gpu_steps
    t1 = time.perf_counter()
    for i in range(n_gpu):
        st = eng.step(last + 1 + i, shift, dtau)
    gpu_step = (time.perf_counter() - t1) / n_gpu
    times = eng.kernel_times()
    sim.close()
    # the oracle's initiator step on one core from the same vector, bounded in time
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    model = O.mom1d(M, N, 1.0, 1.0)
    sty = O.style("dynamic")
    k, v = keys, vals
    n_cpu = 0
    t2 = time.perf_counter()
    while n_cpu < 2000 and time.perf_counter() - t2 < args.init_cpu_seconds:
        k, v, _ = O.step_initiator(model, sty, k, v, rule="initiator", threshold=1.0, seed=17,
                                   step=last + 1 + n_cpu, shift=shift, dtau=dtau, order=0)
        n_cpu += 1
    cpu_step = (time.perf_counter() - t2) / max(n_cpu, 1)
    return {"workload": "HubbardMom1D M=20 N=10 (10 bosons in mode 10) u=1, IsDynamicSemistochastic, "
                        "Initiator(1), ProjectedEnergy, DoubleLogUpdate target 40000, dtau 1e-4",
            "steps": last, "seconds": run_s, "steps_per_s": last / run_s,
            "walkers": float(df["norm"].to_numpy()[-1]), "len": int(df["len"].to_numpy()[-1]),
            "shift_mean_second_half": e_shift, "projected_energy_second_half": e_proj,
            "gpu_ms_per_step_steady": 1e3 * gpu_step,
            "gpu_kernels_ms_per_step": {n: ms / n_gpu for n, (ms, c) in times.items()},
            "gpu_kernel_sum_ms_per_step": sum(ms for ms, c in times.values()) / n_gpu,
            "cpu_baseline": {"ms_per_step": 1e3 * cpu_step, "cores": 1, "kind": "port",
                             "sample": f"{n_cpu} oracle initiator steps from the final vector "
                                       f"({len(keys)} entries), order of the reference's columns"},
            "cpu_over_gpu_steady": cpu_step / gpu_step}


def run_initiator_headline(rimu, keys, vals, shift, step0, args):
    """The equilibrated 1e8-walker HubbardReal1D 20x20 state stepped with Initiator(1)."""
    M = N = 20
    H = rimu.HubbardReal1D(rimu.near_uniform(N, M), u=6.0, t=1.0)
    style = rimu.IsDynamicSemistochastic()
    target = args.target
    eng = rimu._native.Engine(H.descriptor(), style.descriptor(), 17, int(round(2 * target + 100)) + 1,
                              int(min(4 * target + 4096, 2 ** 32 - 2)))
    eng.set_initiator(rimu._native.INITIATOR_DEFAULT, 1.0)
    eng.upload(keys, vals)
    dtau = 1e-3
    for i in range(3):
        eng.step(step0 + i, shift, dtau)
    eng.set_timing(True)
    K = args.init_steps
    S_tot = L_tot = Lnew_tot = 0.0
    L_prev = len(eng)
    t0 = time.perf_counter()
    for i in range(K):
        st = eng.step(step0 + 3 + i, shift, dtau)
        b, records = eng.last_traffic()
        S_tot += records - L_prev
        L_tot += L_prev
        Lnew_tot += st.len
        L_prev = st.len
    elapsed = time.perf_counter() - t0
    times = eng.kernel_times()
    eng.close()
    S, L, Ln = S_tot / K, L_tot / K, Lnew_tot / K
    kern = {}
    for name, (ms, c) in times.items():
        kern[name] = {"ms_avg": ms / max(c, 1), "launches_per_step": c / K}
        bpl = kernel_alg_bytes(name, S, L, Ln)
        if name == "merge":
            bpl += L * 8   # the old values, read for the diagonal deposit's class
        if bpl is not None:
            kern[name]["alg_GBps"] = bpl / (kern[name]["ms_avg"] / 1e3) / 1e9
    return {"workload": "HubbardReal1D M=N=20 u=6 equilibrated plain state (target 1e8), "
                        "stepped with Initiator(1), dtau 1e-3, fixed shift",
            "ms_per_step": 1e3 * elapsed / K, "steps": K, "len": st.len, "walkers": st.norm,
            "spawn_records_per_step": S, "kernels": kern}


def run_wide_case(rimu, args):
    """benchmark/benchmarks.jl:66-73, "(50, 50) Real space" (99-bit addresses, two chunks)
    through lomc, then steady-state step times of the engine and of the oracle."""
    M = N = 50
    addr = rimu.near_uniform(N, M)
    H = rimu.HubbardReal1D(addr, u=6.0)
    dv = rimu.PDVec((addr, 1.0), style=rimu.IsDynamicSemistochastic())
    s_strat = rimu.DoubleLogUpdate(target_walkers=50_000)
    dtau, last = 1e-4, args.wide_laststep
    t0 = time.perf_counter()
    df, sim = rimu.lomc(H, dv, s_strat=s_strat, dtau=dtau, laststep=last, random_seed=17)
    run_s = time.perf_counter() - t0
    shift = float(df["shift"].to_numpy()[-1])
    keys, vals = sim.engine.download()
    eng = sim.engine
    n_gpu = args.wide_gpu_steps
    # wall time per step without kernel timing (one host synchronisation per step), then the
    # same number of steps again with per-kernel events for the kernel-time breakdown
    t1 = time.perf_counter()
    for i in range(n_gpu):
        eng.step(last + 1 + i, shift, dtau)
    gpu_step = (time.perf_counter() - t1) / n_gpu
    eng.set_timing(True)
    for i in range(n_gpu):
        eng.step(last + 1 + n_gpu + i, shift, dtau)
    times = eng.kernel_times()
    sim.close()
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    model = O.real1d(M, N, 6.0, 1.0)
    sty = O.style("dynamic")
    k, v = keys, vals
    n_cpu = 0
    t2 = time.perf_counter()
    while n_cpu < 2000 and time.perf_counter() - t2 < args.wide_cpu_seconds:
        k, v, _ = O.step(model, sty, k, v, seed=17, step=last + 1 + n_cpu, shift=shift,
                         dtau=dtau, order=0)
        n_cpu += 1
    cpu_step = (time.perf_counter() - t2) / max(n_cpu, 1)
    return {"workload": "HubbardReal1D M=N=50 near_uniform u=6, PDVec(addr => 1.0) "
                        "IsDynamicSemistochastic, DoubleLogUpdate target 50000, dtau 1e-4",
            "steps": last, "seconds": run_s, "steps_per_s": last / run_s,
            "walkers": float(df["norm"].to_numpy()[-1]), "len": int(df["len"].to_numpy()[-1]),
            "shift_last": shift,
            "gpu_ms_per_step_steady": 1e3 * gpu_step,
            "gpu_kernels_ms_per_step": {n: ms / n_gpu for n, (ms, c) in times.items()},
            "gpu_kernel_sum_ms_per_step": sum(ms for ms, c in times.values()) / n_gpu,
            "cpu_baseline": {"ms_per_step": 1e3 * cpu_step, "cores": 1, "kind": "port",
                             "sample": f"{n_cpu} oracle steps from the final vector "
                                       f"({len(keys)} entries), order of the reference's columns"},
            "cpu_over_gpu_steady": cpu_step / gpu_step}


def run_wide_headline(rimu, args):
    """HubbardReal1D M = N = 50 (two-chunk addresses) grown to args.wide_target walkers with
    the headline's settings: spawn attempts/s and per-kernel times with 16-byte keys."""
    M = N = 50
    H = rimu.HubbardReal1D(rimu.near_uniform(N, M), u=6.0, t=1.0)
    style = rimu.IsDynamicSemistochastic()
    target = args.wide_target
    eng = rimu._native.Engine(H.descriptor(), style.descriptor(), 17, int(round(2 * target + 100)) + 1,
                              int(min(4 * target + 4096, 2 ** 32 - 2)))
    eng.set_radix_mode(args.radix_mode)
    start = H.starting_address()
    eng.upload(rimu.addresses.pack_keys([start.bits], 2), np.array([10.0]))
    dtau = 1e-3
    strat = rimu.DoubleLogUpdate(target_walkers=target)
    sp = rimu.strategies.DefaultShiftParameters(H.diagonal_element(start), 10.0, dtau)
    step = 0
    t_eq = time.perf_counter()
    last_log = t_eq
    while step < args.wide_equilibration:
        step += 1
        st = eng.step(step, sp.shift, dtau)
        strat.update(sp, st.norm)
        if time.perf_counter() - last_log > 10:
            log(f"wide equilibration step {step}: len {st.len:.3g} norm {st.norm:.3g} shift {sp.shift:.4f}")
            last_log = time.perf_counter()
    t_eq = time.perf_counter() - t_eq
    eng.set_timing(True)
    K = args.wide_steps
    attempts = S_tot = L_tot = Lnew_tot = 0.0
    L_prev = len(eng)
    t0 = time.perf_counter()
    for _ in range(K):
        step += 1
        st = eng.step(step, sp.shift, dtau)
        strat.update(sp, st.norm)
        attempts += st.spawn_attempts
        b, records = eng.last_traffic()
        S_tot += records - L_prev
        L_tot += L_prev
        Lnew_tot += st.len
        L_prev = st.len
    elapsed = time.perf_counter() - t0
    times = eng.kernel_times()
    eng.close()
    S, L, Ln = S_tot / K, L_tot / K, Lnew_tot / K
    kern = {}
    for name, (ms, c) in times.items():
        kern[name] = {"ms_avg": ms / max(c, 1), "launches_per_step": c / K}
        bpl = kernel_alg_bytes(name, S, L, Ln, ksize=16)
        if bpl is not None:
            kern[name]["alg_GBps"] = bpl / (kern[name]["ms_avg"] / 1e3) / 1e9
    return {"workload": f"HubbardReal1D M=N=50 u=6 t=1 (99-bit addresses), IsDynamicSemistochastic "
                        f"FP64, DoubleLogUpdate target {target:.3g}, dtau 1e-3, seed 17",
            "equilibration_steps": args.wide_equilibration, "equilibration_s": t_eq,
            "value": attempts / elapsed, "unit": "spawn_attempts/s",
            "ms_per_step": 1e3 * elapsed / K, "steps": K, "len": st.len, "walkers": st.norm,
            "spawn_records_per_step": S, "kernels": kern}


def run_realspace_leg(rimu, args):
    """HubbardRealSpace: (a) 2D Bose-Hubbard on PeriodicBoundaries(8, 8) with 64 bosons,
    u = 6 (127-bit addresses) grown to args.rs_target walkers, spawn attempts/s and
    per-kernel times, the oracle's step on a bounded sample beside it; (b) the reference's
    "2D Hubbard" system (benchmark/benchmarks.jl:8-14: FermiFS2C(16, 1 => 1, 2 => 1,
    1 => -1, 2 => -1) on PeriodicBoundaries(4, 4)) projected to its ground state with
    IsDeterministic."""
    out = {}
    # (a)
    H = rimu.HubbardRealSpace(rimu.near_uniform(64, 64), geometry=rimu.PeriodicBoundaries(8, 8), u=[[6.0]])
    style = rimu.IsDynamicSemistochastic()
    target = args.rs_target
    eng = rimu._native.Engine(H.descriptor(), style.descriptor(), 17, int(round(2 * target + 100)) + 1,
                              int(min(4 * target + 4096, 2 ** 32 - 2)))
    eng.set_radix_mode(args.radix_mode)
    start = H.starting_address()
    eng.upload(rimu.addresses.pack_keys([start.bits], H.key_words), np.array([10.0]))
    dtau = 1e-3
    strat = rimu.DoubleLogUpdate(target_walkers=target)
    sp = rimu.strategies.DefaultShiftParameters(H.diagonal_element(start), 10.0, dtau)
    step = 0
    t_eq = time.perf_counter()
    last_log = t_eq
    while step < args.rs_equilibration:
        step += 1
        st = eng.step(step, sp.shift, dtau)
        strat.update(sp, st.norm)
        if time.perf_counter() - last_log > 10:
            log(f"realspace equilibration step {step}: len {st.len:.3g} norm {st.norm:.3g} shift {sp.shift:.4f}")
            last_log = time.perf_counter()
    t_eq = time.perf_counter() - t_eq
    eng.set_timing(True)
    K = args.rs_steps
    attempts = S_tot = L_tot = Lnew_tot = 0.0
    L_prev = len(eng)
    t0 = time.perf_counter()
    for _ in range(K):
        step += 1
        st = eng.step(step, sp.shift, dtau)
        strat.update(sp, st.norm)
        attempts += st.spawn_attempts
        b, records = eng.last_traffic()
        S_tot += records - L_prev
        L_tot += L_prev
        Lnew_tot += st.len
        L_prev = st.len
    elapsed = time.perf_counter() - t0
    times = eng.kernel_times()
    keys, vals = eng.download()
    eng.close()
    S, L, Ln = S_tot / K, L_tot / K, Lnew_tot / K
    kern = {}
    for name, (ms, c) in times.items():
        kern[name] = {"ms_avg": ms / max(c, 1), "launches_per_step": c / K}
        bpl = kernel_alg_bytes(name, S, L, Ln, ksize=16)
        if bpl is not None:
            kern[name]["alg_GBps"] = bpl / (kern[name]["ms_avg"] / 1e3) / 1e9
    gpu_step = elapsed / K
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    om = O.real_space((8, 8), [("bose", 64)], u=[[6.0]])
    stride = max(1, len(keys) // 200_000)
    k, v = np.ascontiguousarray(keys[::stride]), np.ascontiguousarray(vals[::stride])
    n_cpu, att_cpu = 0, 0.0
    t2 = time.perf_counter()
    while n_cpu < 20 and time.perf_counter() - t2 < args.rs_cpu_seconds:
        k, v, ost = O.step(om, O.style("dynamic"), k, v, seed=17, step=step + 1 + n_cpu,
                           shift=sp.shift, dtau=dtau, order=0)
        att_cpu += ost["attempts"]
        n_cpu += 1
    cpu_s = time.perf_counter() - t2
    out["bose_8x8"] = {
        "workload": f"HubbardRealSpace BoseFS 64 bosons on PeriodicBoundaries(8, 8), u=6, t=1 "
                    f"(127-bit addresses), IsDynamicSemistochastic FP64, DoubleLogUpdate target {target:.3g}, "
                    f"dtau 1e-3, seed 17",
        "equilibration_steps": args.rs_equilibration, "equilibration_s": t_eq,
        "value": attempts / elapsed, "unit": "spawn_attempts/s", "ms_per_step": 1e3 * gpu_step,
        "steps": K, "len": st.len, "walkers": st.norm, "shift": sp.shift,
        "spawn_records_per_step": S, "kernels": kern,
        "cpu_baseline": {"value": att_cpu / cpu_s, "unit": "spawn_attempts/s", "cores": 1, "kind": "port",
                         "sample": f"{n_cpu} oracle steps from every {stride}-th entry of the final vector "
                                   f"({len(keys[::stride])} entries), {cpu_s:.1f} s"}}
    # (b)
    a = rimu.FermiFS2C(16, (1, 1), (2, 1), (1, -1), (2, -1))
    H2 = rimu.HubbardRealSpace(a, geometry=rimu.PeriodicBoundaries(4, 4))
    dv = rimu.DVec((a, 1.0), style=rimu.IsDeterministic())
    steps = args.rs_hubbard_steps
    t3 = time.perf_counter()
    df, sim = rimu.lomc(H2, dv, laststep=steps, dtau=0.02, s_strat=rimu.DoubleLogUpdate(target_walkers=1e6))
    v = sim.state_vector
    e = rimu.dot(v, H2, v) / rimu.dot(v, v)
    run_s = time.perf_counter() - t3
    out["hubbard_2d_4x4"] = {
        "workload": "HubbardRealSpace(FermiFS2C(16, 1 => 1, 2 => 1, 1 => -1, 2 => -1); "
                    "geometry = PeriodicBoundaries(4, 4)), u = t = 1, IsDeterministic through lomc, dtau 0.02",
        "steps": steps, "seconds": run_s, "steps_per_s": steps / run_s, "len": int(df["len"].to_numpy()[-1]),
        "energy_rayleigh": e, "shift_last": float(df["shift"].to_numpy()[-1])}
    sim.close()
    return out


def run_adjoint_leg(rimu, args):
    """The differentiable projector: HubbardReal1D M = N = 12 (dimension 1,352,078),
    u in {2,4,6,8} as four right-hand sides of one product, K = 2000 steps of dtau 1e-3,
    checkpoints every 45 steps.  grads/s = (dE/du, dE/dt) pairs per second.  The plain
    engine has no adjoint: a user of it differentiates by central differences, four
    K-step IsDeterministic runs per u, one instance per launch; that figure uses the
    plain engine's step time on the full vector."""
    M = N = 12
    us = [2.0, 4.0, 6.0, 8.0]
    K, dtau, ckpt = args.adjoint_steps, 1e-3, 45
    hs = [rimu.HubbardReal1D(rimu.near_uniform(N, M), u=u, t=1.0) for u in us]
    d = rimu.DenseHamiltonians(hs)
    v0 = d.unit_vector(hs[0].starting_address())
    shift = np.zeros(len(us))
    d.projector_energy(20, dtau, shift, v0, ckpt=5)   # warm-up
    t0 = time.perf_counter()
    res = d.projector_energy(K, dtau, shift, v0, ckpt=ckpt)
    elapsed = time.perf_counter() - t0
    grads = len(us) / elapsed
    basis = d.basis()
    vf = res["v"]
    d.close()
    # plain engine, IsDeterministic, on the final vector of u = us[0]
    h = hs[0]
    eng = rimu._native.Engine(h.descriptor(), rimu.IsDeterministic().descriptor(), 1,
                              len(basis) + 1, int(len(basis) * (2 * N + 1) + 4096))
    eng.upload(basis, np.ascontiguousarray(vf[:, 0]))
    for i in range(3):
        eng.step(i + 1, 0.0, dtau)
    t1 = time.perf_counter()
    for i in range(args.plain_sample_steps):
        eng.step(i + 4, 0.0, dtau)
    t_step = (time.perf_counter() - t1) / args.plain_sample_steps
    eng.close()
    plain_s = len(us) * 4 * K * t_step
    # one of those runs end to end (from the start address, K steps) against K x t_step
    eng = rimu._native.Engine(h.descriptor(), rimu.IsDeterministic().descriptor(), 1,
                              len(basis) + 1, int(len(basis) * (2 * N + 1) + 4096))
    eng.upload(np.array([h.starting_address().bits], np.uint64), np.array([1.0]))
    t2 = time.perf_counter()
    for i in range(K):
        eng.step(i + 1, 0.0, dtau)
    e2e = time.perf_counter() - t2
    eng.close()
    return {"metric": "grads/s", "value": grads, "plain_value": len(us) / plain_s,
            "speedup": grads / (len(us) / plain_s), "seconds": elapsed, "dim": len(basis),
            "K": K, "checkpoint_every": ckpt, "u": us,
            "energy": res["energy"].tolist(), "dE_du": res["grad_u"].tolist(),
            "dE_dt": res["grad_t"].tolist(),
            "plain_estimate": (f"central differences on the plain engine: 4 runs x {K} "
                               f"IsDeterministic steps per u at {1e3 * t_step:.3f} ms/step "
                               f"(measured on the full vector), {plain_s:.1f} s"),
            "plain_end_to_end_check": {"measured_s": e2e, "estimated_s": K * t_step,
                                       "measured_over_estimated": e2e / (K * t_step),
                                       "speedup_scaled_to_measured": grads / (len(us) / plain_s) * e2e / (K * t_step)}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--target", type=float, default=1e8)
    ap.add_argument("--equilibration", type=int, default=400,
                    help="steps of population growth and settling before warmup")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-sample-entries", type=int, default=2_000_000)
    ap.add_argument("--cpu-sample-steps", type=int, default=3)
    ap.add_argument("--no-timing", action="store_true", help="disable per-kernel HIP events")
    ap.add_argument("--kernel-steps", type=int, default=10,
                    help="steps after the timed ones with per-kernel HIP events (kernel times, roofline)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy (scale for HBM byte counters)")
    ap.add_argument("--batched-runs", type=int, default=256)
    ap.add_argument("--batched-steps", type=int, default=10_000)
    ap.add_argument("--batched-tune-at", type=int, default=1000)
    ap.add_argument("--adjoint-steps", type=int, default=2000)
    ap.add_argument("--plain-sample-steps", type=int, default=20)
    ap.add_argument("--plain-e2e-runs", type=int, default=8,
                    help="sweep runs also run end to end on the plain engine to check the estimate")
    ap.add_argument("--mom-target", type=float, default=2e7)
    ap.add_argument("--mom-equilibration", type=int, default=12_000,
                    help="cap on growth steps of the momentum-space run")
    ap.add_argument("--mom-settle", type=int, default=200)
    ap.add_argument("--mom-steps", type=int, default=40)
    ap.add_argument("--no-momentum", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--batched-details", default=None, help="write the per-run energy check to this JSON file")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-initiator", action="store_true")
    ap.add_argument("--no-checkpoint", action="store_true")
    ap.add_argument("--no-overlaps", action="store_true")
    ap.add_argument("--overlap-reps", type=int, default=5)
    ap.add_argument("--init-steps", type=int, default=10)
    ap.add_argument("--init-laststep", type=int, default=8000,
                    help="steps of the initiator benchmark case (0 skips it)")
    ap.add_argument("--init-gpu-steps", type=int, default=200)
    ap.add_argument("--init-cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-wide", action="store_true")
    ap.add_argument("--wide-laststep", type=int, default=1000,
                    help="steps of the (50, 50) real-space case (0 skips it)")
    ap.add_argument("--wide-gpu-steps", type=int, default=200)
    ap.add_argument("--wide-cpu-seconds", type=float, default=10.0)
    ap.add_argument("--wide-target", type=float, default=1e8)
    ap.add_argument("--wide-equilibration", type=int, default=800)
    ap.add_argument("--wide-steps", type=int, default=10)
    ap.add_argument("--no-realspace", action="store_true")
    ap.add_argument("--rs-target", type=float, default=1e8)
    ap.add_argument("--rs-equilibration", type=int, default=700)
    ap.add_argument("--rs-steps", type=int, default=10)
    ap.add_argument("--rs-cpu-seconds", type=float, default=10.0)
    ap.add_argument("--rs-hubbard-steps", type=int, default=2000)
    ap.add_argument("--radix-mode", type=int, default=2, choices=(0, 1, 2),
                    help="spawn record sort and merge: 2 top-bit sort and bucket merge, 1 look-back "
                         "scatters over every bit and the merge path, 0 per-pass digit counts")
    ap.add_argument("--state-file", default=None,
                    help="resume from this equilibrated state if it exists, else save it there")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    if args.copy_calibration:
        import torch
        n = 2 ** 27  # 2^27 doubles = 1 GiB read + 1 GiB written per copy
        x = torch.ones(n, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 8 * n, "copy_bytes_written": 8 * n, "copies": 5}))
        return

    rimu = load_package()
    if rimu.device_count() < 1:
        raise SystemExit("bench.py needs a GPU")
    M = N = 20
    H = rimu.HubbardReal1D(rimu.near_uniform(N, M), u=6.0, t=1.0)
    style = rimu.IsDynamicSemistochastic()
    dtau = 1e-3
    target = args.target
    max_length = int(round(2 * target + 100))
    eng = rimu._native.Engine(H.descriptor(), style.descriptor(), 17, max_length + 1,
                              int(min(4 * target + 4096, 2 ** 32 - 2)))
    eng.set_radix_mode(args.radix_mode)
    start = H.starting_address()
    eng.upload(np.array([start.bits], dtype=np.uint64), np.array([10.0]))
    strat = rimu.DoubleLogUpdate(target_walkers=target)
    sp = rimu.strategies.DefaultShiftParameters(H.diagonal_element(start), 10.0, dtau)

    step = 0
    resumed = bool(args.state_file) and os.path.exists(args.state_file)
    st = None
    if resumed:
        z = np.load(args.state_file)
        eng.upload(z["keys"], z["vals"])
        sp.shift, sp.pnorm, step = float(z["shift"]), float(z["pnorm"]), int(z["step"])
        log(f"resumed equilibrated state at step {step}: len {len(z['keys'])} shift {sp.shift:.4f}")
    # DoubleLogUpdate runs on the device (the step's shift argument is ignored); sp mirrors
    # the device's shift parameters after every step
    eng.set_shift_update(rimu._native.SHIFT_DOUBLE_LOG, sp.shift, sp.pnorm, target, strat.zeta,
                         strat.xi_value)
    if not resumed:
        t_eq = time.perf_counter()
        last_log = t_eq
        while step < args.equilibration:
            step += 1
            st = eng.step(step, sp.shift, dtau)
            sp.shift, sp.pnorm = eng.shift_parameters()
            if time.perf_counter() - last_log > 10:
                log(f"equilibration step {step}: len {st.len:.3g} norm {st.norm:.3g} shift {sp.shift:.4f}")
                last_log = time.perf_counter()
        t_eq = time.perf_counter() - t_eq
        log(f"equilibrated after {step} steps in {t_eq:.1f} s: len {st.len:.4g} norm {st.norm:.4g} shift {sp.shift:.4f}")
        if args.state_file:
            k, v = eng.download()
            with open(args.state_file, "wb") as f:
                np.savez(f, keys=k, vals=v, shift=sp.shift, pnorm=sp.pnorm, step=step)

    for _ in range(args.warmup):
        step += 1
        st = eng.step(step, sp.shift, dtau)
        sp.shift, sp.pnorm = eng.shift_parameters()

    # the timed steps run without per-kernel events (`value`, `ms_per_step`); the kernel
    # breakdown and the roofline come from the steps after them, with events on the stream
    eng.set_timing(False)
    attempts = 0.0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step += 1
        st = eng.step(step, sp.shift, dtau)
        sp.shift, sp.pnorm = eng.shift_parameters()
        attempts += st.spawn_attempts
    elapsed = time.perf_counter() - t0
    K = args.steps
    value = attempts / elapsed
    ms_per_step = 1e3 * elapsed / K

    times = {}
    S_tot = L_tot = Lnew_tot = 0.0
    alg_bytes = 0.0
    KT = 0 if args.no_timing else args.kernel_steps
    if KT > 0:
        eng.set_timing(True)
        L_prev = len(eng)
        for _ in range(KT):
            step += 1
            st = eng.step(step, sp.shift, dtau)
            sp.shift, sp.pnorm = eng.shift_parameters()
            b, records = eng.last_traffic()
            alg_bytes += b
            S_tot += records - L_prev
            L_tot += L_prev
            Lnew_tot += st.len
            L_prev = st.len
        times = eng.kernel_times()
        eng.set_timing(False)

    roofline = None
    kernels = {}
    if times:
        roofline, kernels = kernel_rooflines(times, S_tot / KT, L_tot / KT, Lnew_tot / KT, KT,
                                             MEASURED_TRAFFIC_RATIO if args.target == 1e8 else None)
        if roofline is not None:
            roofline["traffic_source"] = (
                "algorithmic bytes x the FETCH_SIZE + WRITE_SIZE ratio measured with rocprofv3 on "
                "this workload (profiles/hbm_traffic_real20_dynamic_1e8.md)"
                if roofline["traffic"] is not None else None)

    checkpoint = None
    if not args.no_checkpoint:
        checkpoint = run_checkpoint_leg(rimu, eng, M, N, step, sp.shift)
        log(f"checkpoint: save {checkpoint['save_GBps']:.2f} GB/s, load {checkpoint['load_GBps']:.2f} GB/s, "
            f"file {checkpoint['file_bytes'] / 1e9:.3f} GB for {checkpoint['state_bytes'] / 1e9:.3f} GB")

    overlaps = None
    if not args.no_overlaps:
        overlaps = run_overlap_leg(rimu, eng, H, step, sp.shift, args)
        log("overlaps: " + ", ".join(f"{k} {v['kernel_ms']:.3f} ms" for k, v in overlaps.items()
                                     if isinstance(v, dict)))

    cpu = None
    if not args.no_cpu:
        keys, vals = eng.download()
        cpu = run_cpu_baseline(keys, vals, args, step + 1, sp.shift)
        log(f"cpu baseline: {cpu['value']:.4g} attempts/s on {cpu['cores']} threads ({cpu['sample']})")

    final_norm, final_len = st.norm, st.len
    init_headline = None
    if not args.no_initiator:
        keys, vals = eng.download()
        eng.close()
        init_headline = run_initiator_headline(rimu, keys, vals, sp.shift, step + 1, args)
        del keys, vals
        log(f"initiator headline: {init_headline['ms_per_step']:.2f} ms/step")
    else:
        eng.close()
    batched = None if args.no_batched else run_batched_leg(rimu, args)
    if batched:
        log(f"batched: {batched['value']:.3g} solves/s vs plain {batched['plain_value']:.3g} "
            f"({batched['speedup']:.1f}x)")
        chk = batched["plain_end_to_end_check"]
        if chk["measured_over_estimated"] is not None:
            log(f"batched: plain end to end on {len(chk['runs'])} runs {chk['measured_s']:.1f} s against "
                f"{chk['estimated_s']:.1f} s estimated ({chk['measured_over_estimated']:.2f}); speedup at the "
                f"measured rate {chk['speedup_scaled_to_measured']:.1f}x")
    momentum = None if args.no_momentum else run_momentum_leg(rimu, args)
    if momentum:
        log(f"momentum: {momentum['value']:.3g} attempts/s, {momentum['ms_per_step']:.2f} ms/step")
    initiator = None
    if not args.no_initiator and args.init_laststep > 0:
        initiator = run_initiator_case(rimu, args)
        log(f"initiator case: {initiator['seconds']:.1f} s for {initiator['steps']} steps, steady "
            f"{initiator['gpu_ms_per_step_steady']:.3f} ms/step GPU vs "
            f"{initiator['cpu_baseline']['ms_per_step']:.1f} ms/step CPU")
    if init_headline is not None:
        initiator = dict(initiator or {}, headline=init_headline)
    wide = None
    if not args.no_wide:
        wide = {}
        if args.wide_laststep > 0:
            wide["case"] = run_wide_case(rimu, args)
            log(f"wide case: {wide['case']['seconds']:.1f} s for {wide['case']['steps']} steps, steady "
                f"{wide['case']['gpu_ms_per_step_steady']:.3f} ms/step GPU "
                f"(kernels {wide['case']['gpu_kernel_sum_ms_per_step']:.3f}) vs "
                f"{wide['case']['cpu_baseline']['ms_per_step']:.1f} ms/step CPU")
        if args.wide_steps > 0:
            wide["headline"] = run_wide_headline(rimu, args)
            log(f"wide headline: {wide['headline']['value']:.3g} attempts/s, "
                f"{wide['headline']['ms_per_step']:.2f} ms/step")
    realspace = None
    if not args.no_realspace:
        realspace = run_realspace_leg(rimu, args)
        log(f"realspace: bose 8x8 {realspace['bose_8x8']['value']:.3g} attempts/s, "
            f"{realspace['bose_8x8']['ms_per_step']:.2f} ms/step; 2D Hubbard 4x4 E = "
            f"{realspace['hubbard_2d_4x4']['energy_rayleigh']:.10f} in {realspace['hubbard_2d_4x4']['seconds']:.2f} s")
    adjoint = None if args.no_adjoint else run_adjoint_leg(rimu, args)
    if adjoint:
        log(f"adjoint: {adjoint['value']:.3g} grads/s vs plain {adjoint['plain_value']:.3g} "
            f"({adjoint['speedup']:.1f}x)")
        chk = adjoint["plain_end_to_end_check"]
        log(f"adjoint: one plain run end to end {chk['measured_s']:.2f} s against {chk['estimated_s']:.2f} s "
            f"estimated ({chk['measured_over_estimated']:.2f}); speedup at the measured rate "
            f"{chk['speedup_scaled_to_measured']:.1f}x")

    out = {
        "metric": "spawned walkers/s & HBM % roofline; batched-U solves/s + grads/s vs plain",
        "value": value,
        "unit": "spawn_attempts/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded FCIQMC run from the near-uniform start address)",
        "config": {"workload": "HubbardReal1D M=N=20 u=6 t=1, IsDynamicSemistochastic FP64, "
                               "target_walkers=1e8, dtau=1e-3, seed 17",
                   "shift_update": "DoubleLogUpdate on the device",
                   "equilibration_steps": args.equilibration, "parallelism": "single GPU"},
        "step_alg_roofline_frac": (alg_bytes / KT) / (ms_per_step / 1e3) / HBM_PEAK if KT else None,
        "walkers": final_norm,
        "len": final_len,
        "spawn_records_per_step": S_tot / KT if KT else None,
        "kernel_steps": KT,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": kernels,
        "batched": batched,
        "adjoint": adjoint,
        "momentum": momentum,
        "initiator": initiator,
        "checkpoint": checkpoint,
        "overlaps": overlaps,
        "wide": wide,
        "realspace": realspace,
    }
    print(json.dumps(_finite(out), allow_nan=False), flush=True)


if __name__ == "__main__":
    main()
