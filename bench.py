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

# HBM bytes per launch measured on this workload (target 1e8 walkers) with rocprofv3
# FETCH_SIZE / WRITE_SIZE, scaled by the same counters on a plain 1 GiB copy (fetch x2.0,
# write x1.0), expressed as a ratio to the kernel's algorithmic bytes.
MEASURED_TRAFFIC_RATIO = {"radix_scatter": 2.462 / 2.283, "merge": 2.290 / 2.295,
                          "spawn": 2.258 / 1.962, "plan": 1.127 / 1.096}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_alg_bytes(name, S, L, Lnew, vsize=8):
    """Algorithmic bytes of one launch: every record the kernel must touch, read or
    written once.  Records are (u64 key, 8-byte value)."""
    rec = 8 + vsize
    if name == "radix_scatter":
        return 2 * S * rec
    if name == "radix_count":
        return S * 8
    if name == "merge":
        return S * rec + L * (8 + vsize) + Lnew * rec
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
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy (scale for HBM byte counters)")
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
    start = H.starting_address()
    eng.upload(np.array([start.bits], dtype=np.uint64), np.array([10.0]))
    strat = rimu.DoubleLogUpdate(target_walkers=target)
    sp = rimu.strategies.DefaultShiftParameters(H.diagonal_element(start), 10.0, dtau)

    step = 0
    if args.state_file and os.path.exists(args.state_file):
        z = np.load(args.state_file)
        eng.upload(z["keys"], z["vals"])
        sp.shift, sp.pnorm, step = float(z["shift"]), float(z["pnorm"]), int(z["step"])
        log(f"resumed equilibrated state at step {step}: len {len(z['keys'])} shift {sp.shift:.4f}")
        st = None
    else:
        t_eq = time.perf_counter()
        last_log = t_eq
        while step < args.equilibration:
            step += 1
            st = eng.step(step, sp.shift, dtau)
            strat.update(sp, st.norm)
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
        strat.update(sp, st.norm)

    eng.set_timing(not args.no_timing)
    attempts = 0.0
    S_tot = L_tot = Lnew_tot = 0.0
    alg_bytes = 0.0
    L_prev = len(eng)
    t0 = time.perf_counter()
    for _ in range(args.steps):
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
    elapsed = time.perf_counter() - t0
    times = eng.kernel_times() if not args.no_timing else {}
    K = args.steps
    value = attempts / elapsed
    ms_per_step = 1e3 * elapsed / K

    roofline = None
    kernels = {}
    if times:
        for name, (ms, launches) in times.items():
            kernels[name] = {"ms_total": ms, "launches": launches, "ms_avg": ms / max(launches, 1)}
        dom = max((n for n in times if kernel_alg_bytes(n, 1, 1, 1) is not None),
                  key=lambda n: times[n][0])
        ms, launches = times[dom]
        per_step_launches = launches / K
        bytes_per_launch = kernel_alg_bytes(dom, S_tot / K, L_tot / K, Lnew_tot / K)
        avg_s = ms / launches / 1e3
        achieved = bytes_per_launch / avg_s
        traffic = None
        if dom in MEASURED_TRAFFIC_RATIO and args.target == 1e8:
            traffic = MEASURED_TRAFFIC_RATIO[dom] * bytes_per_launch
        roofline = {"bound": "hbm", "kernel": dom, "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                    "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": traffic,
                    "alg_bytes_per_launch": bytes_per_launch, "avg_launch_ms": avg_s * 1e3,
                    "launches_per_step": per_step_launches}
        for name in kernels:
            bpl = kernel_alg_bytes(name, S_tot / K, L_tot / K, Lnew_tot / K)
            if bpl is not None:
                kernels[name]["alg_GBps"] = bpl / (kernels[name]["ms_avg"] / 1e3) / 1e9

    cpu = None
    if not args.no_cpu:
        keys, vals = eng.download()
        cpu = run_cpu_baseline(keys, vals, args, step + 1, sp.shift)
        log(f"cpu baseline: {cpu['value']:.4g} attempts/s on {cpu['cores']} threads ({cpu['sample']})")

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
                   "equilibration_steps": args.equilibration, "parallelism": "single GPU"},
        "step_alg_roofline_frac": alg_bytes / elapsed / HBM_PEAK,
        "walkers": st.norm,
        "len": st.len,
        "spawn_records_per_step": S_tot / K,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": kernels,
    }
    eng.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
