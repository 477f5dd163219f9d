"""Shift strategies and post-step strategies.

References: ``src/strategies_and_params/shiftstrategy.jl`` (``DefaultShiftParameters``
:32-38, ``DontUpdate`` :77-91, ``LogUpdateAfterTargetWalkers`` :100-122, ``LogUpdate``
:135-146, ``DoubleLogUpdate`` :161-181, ``DoubleLogUpdateAfterTargetWalkers`` :190-215),
``src/strategies_and_params/poststepstrategy.jl:82-121`` (``ProjectedEnergy``),
``src/strategies_and_params/timestepstrategy.jl:16-22`` (``ConstantTimeStep``).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .dictvectors import ADDRESS_TYPES, DVec, FrozenDVec, freeze
from .addresses import BoseFS, unpack_keys


@dataclass
class DefaultShiftParameters:
    shift: float
    pnorm: float
    time_step: float
    counter: int = 0
    shift_mode: bool = False


def _log(x):
    return math.log(x) if x > 0 else -math.inf


@dataclass(frozen=True)
class DontUpdate:
    target_walkers: float = 1000

    def update(self, sp, tnorm):
        return {"shift": sp.shift, "norm": tnorm}, tnorm < self.target_walkers


@dataclass(frozen=True)
class LogUpdate:
    zeta: float = 0.08

    def update(self, sp, tnorm):
        sp.shift -= self.zeta / sp.time_step * _log(tnorm / sp.pnorm)
        sp.pnorm = tnorm
        return {"shift": sp.shift, "norm": tnorm}, True


@dataclass(frozen=True)
class LogUpdateAfterTargetWalkers:
    target_walkers: float = 1000
    zeta: float = 0.08

    def update(self, sp, tnorm):
        if sp.shift_mode or tnorm > self.target_walkers:
            sp.shift_mode = True
            sp.shift -= self.zeta / sp.time_step * _log(tnorm / sp.pnorm)
        sp.pnorm = tnorm
        return {"shift": sp.shift, "norm": tnorm, "shift_mode": sp.shift_mode}, True


@dataclass(frozen=True)
class DoubleLogUpdate:
    """S <- S - xi/dt ln(N/N_target) - zeta/dt ln(N/N_prev) (shiftstrategy.jl:174-181)."""
    target_walkers: float = 1000
    zeta: float = 0.08
    xi: float | None = None

    @property
    def xi_value(self):
        return self.zeta ** 2 / 4 if self.xi is None else self.xi

    def update(self, sp, tnorm):
        dt = sp.time_step
        sp.shift -= self.xi_value / dt * _log(tnorm / self.target_walkers) + self.zeta / dt * _log(tnorm / sp.pnorm)
        sp.pnorm = tnorm
        return {"shift": sp.shift, "norm": tnorm}, True


@dataclass(frozen=True)
class DoubleLogUpdateAfterTargetWalkers:
    target_walkers: float = 1000
    zeta: float = 0.08
    xi: float | None = None

    def update(self, sp, tnorm):
        if sp.shift_mode or tnorm > self.target_walkers:
            sp.shift_mode = True
            xi = self.zeta ** 2 / 4 if self.xi is None else self.xi
            dt = sp.time_step
            sp.shift -= xi / dt * _log(tnorm / self.target_walkers) + self.zeta / dt * _log(tnorm / sp.pnorm)
        sp.pnorm = tnorm
        return {"shift": sp.shift, "norm": tnorm, "shift_mode": sp.shift_mode}, True


@dataclass(frozen=True)
class ConstantTimeStep:
    pass


class ProjectedEnergy:
    """After every step report ``vproj = dot(projector, v)`` and ``hproj = dot(H' projector, v)``
    (poststepstrategy.jl:82-121).  ``H' * projector`` is formed once with the device
    Hamiltonian (``mul!``, abstractdvec.jl:275-284) and frozen."""

    def __init__(self, hamiltonian, projector, vproj="vproj", hproj="hproj"):
        self.vproj_name = vproj
        self.hproj_name = hproj
        self.hamiltonian = hamiltonian
        if isinstance(projector, ADDRESS_TYPES):
            projector = DVec((projector, 1.0))
        self.vproj = freeze(projector)
        self.hproj = self._compute_hproj(hamiltonian, self.vproj)

    @staticmethod
    def _compute_hproj(h, fp: FrozenDVec) -> FrozenDVec:
        keys = fp.keys_bits()
        coefs = fp.coefs()
        diag, rows, cols, vals = h.column(keys)
        acc = {}
        for k, d, c in zip(unpack_keys(keys), diag.tolist(), coefs.tolist()):
            acc[k] = acc.get(k, 0.0) + d * c
        for r, j, v in zip(unpack_keys(rows), cols.tolist(), vals.tolist()):
            acc[r] = acc.get(r, 0.0) + v * coefs[j]
        return FrozenDVec([(h.address.with_bits(k), v) for k, v in acc.items() if v != 0])

    def names(self):
        return (self.vproj_name, self.hproj_name)
