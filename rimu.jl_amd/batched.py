"""Many independent FCIQMC runs advanced together (C ABI ``rimu_batch_*``): each run's
spawning is split over workgroups by its target walker number; a step is five launches
covering all runs or one persistent launch (``set_persistent``), and steps are captured in
a HIP graph.

Each run is ``lomc!``/``ProjectorMonteCarloProblem`` with ``IsStochasticInteger``
walkers on a ``HubbardReal1D`` (``src/lomc.jl:185-195``): start vector
``start_address => population``, ``DoubleLogUpdate`` shift control
(``src/strategies_and_params/shiftstrategy.jl:174-181``) computed on the device, and a
``ProjectedEnergy`` onto the start address (``poststepstrategy.jl:82-121``).
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np

from . import _native
from .addresses import BoseFS, near_uniform
from .hamiltonians import HubbardReal1D

_p = _native._p
STATUS = {0: "running", 1: "population dead", 2: "max_length reached", 3: "spawn buffer overflow"}


@dataclass
class RunSpec:
    hamiltonian: HubbardReal1D
    target_walkers: float = 1000
    dtau: float = 1e-3
    seed: int = 0
    population: int = 10
    shift: float | None = None       # default: H_aa of the start address (lomc!)
    zeta: float = 0.08
    xi: float | None = None
    max_length: int | None = None    # default 2 * target_walkers + 100

    def descriptor(self) -> _native.BatchRun:
        h = self.hamiltonian
        a = h.starting_address()
        d = _native.BatchRun()
        d.M, d.N, d.u, d.t = a.M, a.N, h.u, h.t
        d.seed = self.seed & (2 ** 64 - 1)
        d.start_address = a.bits
        d.start_population = int(self.population)
        d.shift = h.u * sum(n * (n - 1) for n in a.onr) / 2 if self.shift is None else float(self.shift)
        d.target_walkers = float(self.target_walkers)
        d.zeta = float(self.zeta)
        d.xi = self.zeta ** 2 / 4 if self.xi is None else float(self.xi)
        d.dtau = float(self.dtau)
        d.max_length = int(round(2 * self.target_walkers + 100)) if self.max_length is None else int(self.max_length)
        return d


class BatchedRuns:
    FIELDS = ("shift", "norm", "len", "vproj", "hproj", "spawn_attempts")

    def __init__(self, specs, history_steps: int):
        self.specs = list(specs)
        self.n = len(self.specs)
        arr = (_native.BatchRun * self.n)(*[s.descriptor() for s in self.specs])
        self._h = ctypes.c_void_p()
        self.history_steps = int(history_steps)
        _native.check(_native.lib().rimu_batch_create(arr, self.n, self.history_steps, ctypes.byref(self._h)))

    def close(self):
        if self._h:
            _native.lib().rimu_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_config(self, block_size: int, radix_bits: int, persistent: bool | None = None):
        _native.check(_native.lib().rimu_batch_set_config(self._h, block_size, radix_bits))
        if persistent is not None:
            self.set_persistent(persistent)

    def set_persistent(self, persistent: bool):
        """One persistent launch per step (True) or five launches (False); same results."""
        _native.check(_native.lib().rimu_batch_set_persistent(self._h, int(bool(persistent))))

    def set_wait(self, spin_limit: int, withhold_first: bool = False):
        """Bound of the persistent step's waits and the host-forced give-up test
        (``rimu_batch_set_wait``)."""
        _native.check(_native.lib().rimu_batch_set_wait(self._h, int(spin_limit), int(bool(withhold_first))))

    def step(self, nsteps: int, steps_per_graph: int = 100):
        _native.check(_native.lib().rimu_batch_step(self._h, int(nsteps), int(steps_per_graph)))

    def autotune(self, trial_steps: int = 20):
        """Time every (persistent, block, radix) candidate; the batch keeps the fastest.
        Returns (block, radix, persistent, {(persistent, block, radix): ms per step})."""
        ms = np.zeros(18)
        bb, br, bp = ctypes.c_int32(0), ctypes.c_int32(0), ctypes.c_int32(0)
        _native.check(_native.lib().rimu_batch_autotune(self._h, int(trial_steps), ctypes.byref(bb),
                                                        ctypes.byref(br), ctypes.byref(bp), _p(ms)))
        table = {(bool(p), blk, rad): ms[(p * 3 + i) * 3 + j] for p in (0, 1)
                 for i, blk in enumerate((256, 512, 1024)) for j, rad in enumerate((4, 6, 8))}
        return int(bb.value), int(br.value), bool(bp.value), table

    def history(self):
        n = ctypes.c_int64(0)
        bufs = [np.zeros((self.history_steps, self.n)) for _ in self.FIELDS]
        _native.check(_native.lib().rimu_batch_history(self._h, ctypes.byref(n), *[_p(b) for b in bufs]))
        return {f: b[:n.value] for f, b in zip(self.FIELDS, bufs)}

    def status(self):
        st = np.zeros(self.n, dtype=np.int32)
        ln = np.zeros(self.n, dtype=np.int64)
        _native.check(_native.lib().rimu_batch_status(self._h, _p(st), _p(ln)))
        return st, ln

    def vector(self, run: int):
        _, ln = self.status()
        cap = max(int(ln[run]), 1)
        k = np.zeros(cap, dtype=np.uint64)
        v = np.zeros(cap, dtype=np.int64)
        n = ctypes.c_int64(0)
        _native.check(_native.lib().rimu_batch_vector(self._h, int(run), _p(k), _p(v), cap, ctypes.byref(n)))
        return k[:n.value], v[:n.value]


def u_sweep_specs(nruns: int = 256, dtau: float = 1e-3, rng_seed: int = 2024):
    """The batched U-sweep workload: M ~ U{6..14}, N ~ U{ceil(M/2)..M}, u ~ U[1, 10], t = 1,
    target walkers log-uniform in [1e3, 1e5], seeds 0..nruns-1, near-uniform start."""
    rng = np.random.default_rng(rng_seed)
    specs = []
    for i in range(nruns):
        M = int(rng.integers(6, 15))
        N = int(rng.integers(math.ceil(M / 2), M + 1))
        u = float(rng.uniform(1.0, 10.0))
        tw = float(10 ** rng.uniform(3.0, 5.0))
        h = HubbardReal1D(near_uniform(N, M), u=u, t=1.0)
        specs.append(RunSpec(h, target_walkers=tw, dtau=dtau, seed=i))
    return specs


def ground_state_energies(specs):
    """Exact ground-state energy of every run's Hamiltonian: Lanczos on the device over the
    ranked basis (``rimu_dense_ground_state``), runs of equal M, N grouped eight to a call."""
    from .dense import DenseHamiltonians
    e0 = np.full(len(specs), np.nan)
    groups = {}
    for r, sp in enumerate(specs):
        a = sp.hamiltonian.starting_address()
        groups.setdefault((a.M, a.N), []).append(r)
    for _, runs in groups.items():
        for c in range(0, len(runs), 8):
            chunk = runs[c:c + 8]
            d = DenseHamiltonians([specs[r].hamiltonian for r in chunk])
            e, _ = d.ground_state()
            e0[chunk] = e
            d.close()
    return e0


def energy_check(specs, hist, status, e0, skip):
    """Per run: the shift estimator (blocking analysis) and the projected energy
    (``ratio_of_means`` with Monte Carlo error propagation) after ``skip`` steps, and whether
    each holds the exact energy ``e0`` within 2 and 3 standard errors (``val_and_errs``
    intervals, StatsTools/convenience.jl:61-112).  Runs that stopped are left out."""
    from . import statstools as ST
    rows = []
    for r, sp in enumerate(specs):
        if status[r] != 0:
            continue
        se = ST.blocking_analysis(hist["shift"][:, r], skip=skip)
        pe = ST.ratio_of_means(hist["hproj"][:, r], hist["vproj"][:, r], skip=skip)
        a = sp.hamiltonian.starting_address()
        ok_s = se.k > 0 and np.isfinite(se.err)
        rows.append({
            "run": r, "M": a.M, "N": a.N, "u": sp.hamiltonian.u, "target_walkers": sp.target_walkers,
            "mean_walkers": float(hist["norm"][skip:, r].mean()), "E0": float(e0[r]),
            "shift": se.mean, "shift_err": se.err, "shift_k": se.k,
            "shift_2sigma": bool(ok_s and ST.within(se, e0[r], n=2)),
            "shift_3sigma": bool(ok_s and ST.within(se, e0[r], n=3)),
            "projected": pe.pmedian(), "projected_f": pe.f, "projected_sigma_f": pe.sigma_f,
            "projected_err": pe.err, "delta_y": pe.delta_y, "projected_k": pe.k,
            "projected_success": pe.success,
            "projected_2sigma": bool(ST.within(pe, e0[r], n=2)),
            "projected_3sigma": bool(ST.within(pe, e0[r], n=3)),
        })
    return rows
