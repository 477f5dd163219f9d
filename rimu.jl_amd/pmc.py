"""``ProjectorMonteCarloProblem`` / ``init`` / ``solve`` / ``step!`` / ``lomc!`` over the
device engine.

The host keeps Rimu's step loop (``advance!``, ``src/fciqmc.jl:126-181``; ``step!`` and
``solve!``, ``src/pmc_simulation.jl:265-310``, ``:350-452``) and the report; one
``rimu_step`` call replaces ``apply_operator!`` + ``walkernumber_and_length`` + the
ProjectedEnergy post-step.  DoubleLogUpdate (the default shift strategy) runs on the
device after every step; the other shift strategies update on the host.  Keyword defaults
follow ``src/projector_monte_carlo_problem.jl:150-290`` and ``src/lomc.jl:95-195``.
"""
from __future__ import annotations

import math
import secrets
import time

import numpy as np

from . import _native
from .addresses import BoseFS, key_words, pack_keys, unpack_keys
from .dictvectors import ADDRESS_TYPES, DVec, DeviceDVec, FrozenDVec, PDVec
from .hamiltonians import AbstractHamiltonian
from .initiators import Initiator, InitiatorRule, NonInitiator
from .replicas import NoStats, ReplicaStrategy
from .strategies import (ConstantTimeStep, DefaultShiftParameters, DoubleLogUpdate,
                         ProjectedEnergy)
from .styles import IsDynamicSemistochastic, IsStochasticInteger, StochasticStyle

_seed_counter = [0x5EED_0000_0000]


class ProjectorMonteCarloProblem:
    def __init__(self, hamiltonian: AbstractHamiltonian, *, start_at=None, shift=None,
                 style: StochasticStyle | None = None, initiator=False, threading=None,
                 time_step: float = 0.01, starting_step: int = 0, last_step: int = 100,
                 wall_time: float = math.inf, target_walkers: float = 1000, zeta: float = 0.08,
                 xi: float | None = None, shift_strategy=None, time_step_strategy=None,
                 post_step_strategy=(), max_length: int | None = None, metadata=None,
                 display_name: str = "PMCSimulation", random_seed=True, n_replicas: int = 1,
                 replica_strategy: ReplicaStrategy | None = None, reporting_interval: int = 1):
        # projector_monte_carlo_problem.jl:221-223
        if initiator is None or initiator is False:
            initiator = NonInitiator()
        elif initiator is True:
            initiator = Initiator()
        elif not isinstance(initiator, InitiatorRule):
            raise ValueError(f"Invalid initiator {initiator!r}")
        self.initiator = initiator
        # projector_monte_carlo_problem.jl:164, :205: the strategy fixes the replica count
        if replica_strategy is None:
            replica_strategy = NoStats(n_replicas)
        self.replica_strategy = replica_strategy
        self.n_replicas = replica_strategy.n_replicas
        if self.n_replicas < 1:
            raise ValueError("at least one replica")
        self.hamiltonian = hamiltonian
        self.start_at = hamiltonian.starting_address() if start_at is None else start_at
        self.style = IsDynamicSemistochastic() if style is None else style
        self.time_step = float(time_step)
        self.shift = shift
        self.starting_step = int(starting_step)
        self.last_step = int(last_step)
        self.wall_time = float(wall_time)
        if shift_strategy is None:
            shift_strategy = DoubleLogUpdate(target_walkers=target_walkers, zeta=zeta, xi=xi)
        self.shift_strategy = shift_strategy
        self.time_step_strategy = time_step_strategy or ConstantTimeStep()
        if isinstance(post_step_strategy, ProjectedEnergy):
            post_step_strategy = (post_step_strategy,)
        self.post_step_strategy = tuple(post_step_strategy)
        if max_length is None:
            tw = getattr(shift_strategy, "target_walkers", target_walkers)
            max_length = round(2 * abs(tw) + 100)
        self.max_length = int(max_length)
        self.metadata = dict(metadata or {})
        self.metadata["display_name"] = display_name
        if random_seed is True:
            random_seed = secrets.randbits(64)
        elif random_seed is False:
            random_seed = None
        self.random_seed = None if random_seed is None else int(random_seed) & (2 ** 64 - 1)
        self.reporting_interval = int(reporting_interval)

    def __repr__(self):
        return (f"ProjectorMonteCarloProblem(hamiltonian={self.hamiltonian}, style={self.style}, "
                f"time_step={self.time_step}, last_step={self.last_step}, "
                f"shift_strategy={self.shift_strategy}, max_length={self.max_length})")


def default_starting_vector(address: BoseFS, style=None, population=10,
                            initiator: InitiatorRule | None = None) -> DVec:
    """``default_starting_vector`` (qmc_states.jl:229-260): ``PDVec(address =>
    population; style, initiator)``."""
    style = style or IsDynamicSemistochastic()
    return PDVec((address, population), style=style, initiator=initiator or NonInitiator())


def _start_vector(problem: ProjectorMonteCarloProblem) -> DVec:
    """``_set_up_starting_vectors`` (pmc_simulation.jl:33-85): an address or a frozen
    vector gets the problem's style and initiator rule; a ``DVec`` is copied with its own
    style and rule."""
    s = problem.start_at
    if isinstance(s, ADDRESS_TYPES):
        return default_starting_vector(s, problem.style, initiator=problem.initiator)
    if isinstance(s, FrozenDVec):
        return PDVec(list(s.pairs), style=problem.style, initiator=problem.initiator)
    if isinstance(s, DVec):
        return s.copy()
    raise TypeError("start_at must be an address, a FrozenDVec or a DVec")


def rayleigh_quotient_dvec(h: AbstractHamiltonian, v: DVec) -> float:
    """dot(v, H, v) / dot(v, v) (fciqmc.jl:51-61, abstractdvec.jl:313-324)."""
    items = list(v.pairs())
    bits = [k.bits for k, _ in items]
    keys = pack_keys(bits, h.key_words)
    vals = np.array([float(x) for _, x in items])
    diag, rows, cols, offv = h.column(keys)
    index = {k: i for i, k in enumerate(bits)}
    num = float(np.dot(vals * diag, vals))
    for r, c, x in zip(unpack_keys(rows), cols.tolist(), offv.tolist()):
        i = index.get(r)
        if i is not None:
            num += vals[i] * x * vals[c]
    return num / float(np.dot(vals, vals))


class _Replica:
    """One replica: its engine, device vector and shift parameters (SingleState,
    pmc_simulation.jl:126-146).  ``id`` is the report column suffix: "" for a single replica,
    "_i" otherwise."""

    def __init__(self, problem, h, v0, seed, index, n):
        tw = abs(getattr(problem.shift_strategy, "target_walkers", 1000))
        max_entries = problem.max_length + 1
        max_spawns = int(min(max(4 * tw + 4096, 4 * len(v0) + 4096), 2 ** 32 - 2))
        self.style = v0.style
        self.engine = _native.Engine(h.descriptor(), self.style.descriptor(), seed,
                                     max_entries, max_spawns)
        self.initiator = v0.initiator
        self.engine.set_initiator(*self.initiator.descriptor())
        keys, vals = v0.to_arrays()
        self.engine.upload(keys, vals)
        self.v = DeviceDVec(self.engine, h, self.style, self.initiator)
        self.id = "" if n == 1 else f"_{index + 1}"
        self.seed = seed
        self.last_stats = None


class PMCSimulation:
    """State, report and status of a run (pmc_simulation.jl:22-31)."""

    def __init__(self, problem: ProjectorMonteCarloProblem):
        self.problem = problem
        h = problem.hamiltonian
        v0 = _start_vector(problem)
        if problem.random_seed is not None:
            seed = problem.random_seed
        else:
            _seed_counter[0] += 0x9E3779B97F4A7C15
            seed = _seed_counter[0] & (2 ** 64 - 1)
        self.seed = seed
        n = problem.n_replicas
        # replica i draws from its own Philox key; replica 1 keeps the problem's seed
        self.replicas = [_Replica(problem, h, v0, (seed + i * 0xD1B54A32D192ED03) & (2 ** 64 - 1), i, n)
                         for i in range(n)]
        self.style = v0.style
        self.initiator = v0.initiator
        # initial shift: Rayleigh quotient of the start vector unless given
        shift = problem.shift
        if shift is None:
            shift = rayleigh_quotient_dvec(h, v0)
        for r in self.replicas:
            r.shift_parameters = DefaultShiftParameters(float(shift), v0.walkernumber(),
                                                        problem.time_step)
        # DoubleLogUpdate runs on the device (update_shift_parameters!, shiftstrategy.jl:174-181,
        # after every step from the step's norm); the other strategies update on the host
        s = problem.shift_strategy
        self.device_shift = (type(s) is DoubleLogUpdate and s.target_walkers > 0
                             and v0.walkernumber() > 0)
        if self.device_shift:
            for r in self.replicas:
                self._set_device_shift(r)
        # projections (only ProjectedEnergy is evaluated on the device)
        self.projected = [p for p in problem.post_step_strategy if isinstance(p, ProjectedEnergy)]
        if len(self.projected) > 1:
            raise NotImplementedError("one ProjectedEnergy per problem")
        if self.projected:
            p = self.projected[0]
            for r in self.replicas:
                r.engine.set_projector(p.vproj.keys_bits(), p.vproj.coefs(),
                                       p.hproj.keys_bits(), p.hproj.coefs())
        self.step = problem.starting_step
        self.last_step = problem.last_step
        self.wall_time = problem.wall_time
        self.columns = {}
        self.modified = False
        self.aborted = False
        self.success = False
        self.message = ""
        self.elapsed_time = 0.0

    # the first replica's state (the only one for n_replicas = 1)
    @property
    def engine(self):
        return self.replicas[0].engine

    @property
    def v(self) -> DeviceDVec:
        return self.replicas[0].v

    @property
    def shift_parameters(self):
        return self.replicas[0].shift_parameters

    @property
    def last_stats(self):
        return self.replicas[0].last_stats

    @property
    def state_vectors(self):
        return [r.v for r in self.replicas]

    # report ---------------------------------------------------------------------------
    def _report(self, name, value):
        self.columns.setdefault(name, []).append(value)

    @property
    def df(self):
        import pandas as pd
        n = max((len(c) for c in self.columns.values()), default=0)
        cols = {k: v for k, v in self.columns.items() if len(v) == n}
        df = pd.DataFrame(cols)
        df.attrs.update(self.problem.metadata)
        df.attrs.update({"aborted": self.aborted, "success": self.success,
                         "message": self.message, "elapsed_time": self.elapsed_time})
        return df

    def DataFrame(self):
        return self.df

    # stepping -------------------------------------------------------------------------
    def _set_device_shift(self, r: _Replica):
        s, sp = self.problem.shift_strategy, r.shift_parameters
        r.engine.set_shift_update(_native.SHIFT_DOUBLE_LOG, sp.shift, sp.pnorm, s.target_walkers,
                                  s.zeta, s.xi_value)

    def advance(self, r: _Replica | None = None) -> bool:
        """``advance!`` (fciqmc.jl:126-181) of one replica."""
        r = self.replicas[0] if r is None else r
        p = self.problem
        sp = r.shift_parameters
        st = None
        if self.device_shift and (sp.shift, sp.pnorm) != r.engine.shift_parameters():
            self._set_device_shift(r)   # the shift parameters were set from outside
        try:
            st = r.engine.step(self.step, sp.shift, sp.time_step)
        except _native.CapacityError:
            self.message_detail = "`max_length` reached. Aborting."
            return False
        tnorm, length = st.norm, int(st.len)
        r.v._norm = tnorm
        if self.device_shift:
            sp.shift, sp.pnorm = r.engine.shift_parameters()
            shift_stats, proceed = {"shift": sp.shift, "norm": tnorm}, True
        else:
            shift_stats, proceed = p.shift_strategy.update(sp, tnorm)
        if self.step % p.reporting_interval == 0:
            i = r.id
            self._report("len" + i, length)
            for k, val in shift_stats.items():
                self._report(k + i, val)
            for name, val in zip(self.style.step_stat_names(), self.style.stat_values(st)):
                self._report(name + i, val)
            for pe in self.projected:
                self._report(pe.vproj_name + i, st.vproj)
                self._report(pe.hproj_name + i, st.hproj)
        r.last_stats = st
        if length == 0:
            self.message_detail = "Population is dead. Aborting."
            return False
        if length > p.max_length:
            self.message_detail = "`max_length` reached. Aborting."
            return False
        return proceed

    def step_(self):
        """``step!`` (pmc_simulation.jl:265-310)."""
        if self.aborted or self.success or self.step >= self.last_step:
            return self
        self.step += 1
        reporting = self.step % self.problem.reporting_interval == 0
        if reporting:
            self._report("step", self.step)
        proceed = True
        for r in self.replicas:     # every replica advances, as proceed &= advance!(...)
            proceed = self.advance(r) and proceed
        self.modified = True
        if reporting:               # pmc_simulation.jl:293-297
            names, values = self.problem.replica_strategy.replica_stats(self.state_vectors)
            for nm, val in zip(names, values):
                self._report(nm, val)
        if not proceed:
            self.aborted = True
            self.message = f"Aborted in step {self.step}."
            return self
        if self.step == self.last_step:
            self.success = True
        return self

    def solve_(self, last_step=None, wall_time=None):
        """``solve!`` (pmc_simulation.jl:350-452)."""
        if last_step is not None:
            self.last_step = int(last_step)
            self.aborted = self.success = False
        if wall_time is not None:
            self.wall_time = float(wall_time)
            self.aborted = self.success = False
        t0 = time.time() - self.elapsed_time
        while not self.aborted and not self.success:
            if self.step >= self.last_step:
                self.success = True
                break
            if time.time() - t0 > self.wall_time:
                self.aborted = True
                self.message = "Wall time limit reached."
                break
            self.step_()
        self.elapsed_time = time.time() - t0
        return self

    @property
    def state_vector(self) -> DeviceDVec:
        return self.v

    def close(self):
        for r in self.replicas:
            r.engine.close()


def init(problem: ProjectorMonteCarloProblem) -> PMCSimulation:
    return PMCSimulation(problem)


def solve(problem: ProjectorMonteCarloProblem) -> PMCSimulation:
    return init(problem).solve_()


def lomc(ham: AbstractHamiltonian, v=None, *, style=None, targetwalkers=1000, s_strat=None,
         laststep=100, dtau=0.01, shift=None, address=None, post_step=(), maxlength=None,
         step=0, random_seed=None, initiator: InitiatorRule | None = None,
         replica_strategy: ReplicaStrategy | None = None):
    """``lomc!`` (lomc.jl:95-195): default style ``IsStochasticInteger``, default shift the
    diagonal element of the starting address, ``initiator`` the rule of the default start
    vector (unused when ``v`` is given), no re-seeding.  Returns (df, simulation)."""
    address = ham.starting_address() if address is None else address
    if v is None:
        v = default_starting_vector(address, style or IsStochasticInteger(),
                                    initiator=initiator or NonInitiator())
    style = v.style if style is None else style
    if s_strat is None:
        s_strat = DoubleLogUpdate(target_walkers=targetwalkers)
    if shift is None:
        shift = ham.diagonal_element(address)
    tw = getattr(s_strat, "target_walkers", targetwalkers)
    maxlength = round(2 * abs(tw) + 100) if maxlength is None else maxlength
    p = ProjectorMonteCarloProblem(ham, start_at=v, style=style, shift=shift, time_step=dtau,
                                   starting_step=step, last_step=laststep, shift_strategy=s_strat,
                                   post_step_strategy=post_step, max_length=maxlength,
                                   random_seed=False if random_seed is None else random_seed,
                                   replica_strategy=replica_strategy, display_name="lomc!")
    sim = init(p).solve_()
    return sim.df, sim
