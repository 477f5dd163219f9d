"""Stochastic styles and compression strategies (plain data handed to the engine).

References: ``src/StochasticStyles/styles.jl`` (``IsStochasticInteger`` :11-25,
``IsDeterministic`` :76-105, ``IsDynamicSemistochastic`` :175-214, ``step_stats``),
``src/StochasticStyles/compression.jl:8-16`` (``ThresholdCompression``, ``len_before``),
``src/Interfaces/stochasticstyles.jl:88-121`` (``NoCompression``).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

from . import _native


@dataclass(frozen=True)
class NoCompression:
    def step_stat_names(self):
        return ()


@dataclass(frozen=True)
class ThresholdCompression:
    threshold: float = 1.0

    def step_stat_names(self):
        return ("len_before",)


@dataclass(frozen=True)
class WithReplacement:
    threshold: float = 0.0


class StochasticStyle:
    value_type = float
    kind = 0
    compression = NoCompression()

    def step_stat_names(self):
        raise NotImplementedError

    def descriptor(self) -> _native.Style:
        s = _native.Style()
        s.kind = self.kind
        if isinstance(self.compression, ThresholdCompression):
            s.compression = 1
            s.compression_threshold = float(self.compression.threshold)
        else:
            s.compression = 0
            s.compression_threshold = 1.0
        s.proj_threshold = float(getattr(self, "proj_threshold", 0.0))
        s.rel_threshold = float(getattr(self, "rel_spawning_threshold", 1.0))
        s.abs_threshold = float(getattr(self, "abs_spawning_threshold", math.inf))
        return s


@dataclass(frozen=True)
class IsStochasticInteger(StochasticStyle):
    """Integer walkers, stochastic rounding of every deposit (styles.jl:11-25)."""
    value_type = int
    kind = _native.STYLE_INTEGER
    compression = NoCompression()

    def step_stat_names(self):
        return ("spawn_attempts", "spawns", "deaths", "clones", "zombies")

    def stat_values(self, st):
        return (st.spawn_attempts, st.spawns, st.deaths, st.clones, st.zombies)


@dataclass(frozen=True)
class IsDeterministic(StochasticStyle):
    """Exact vector-matrix products (styles.jl:76-105)."""
    compression: object = NoCompression()
    kind = _native.STYLE_DETERMINISTIC

    def step_stat_names(self):
        return ("exact_steps",) + self.compression.step_stat_names()

    def stat_values(self, st):
        base = (st.exact_steps,)
        return base + ((st.len_before,) if isinstance(self.compression, ThresholdCompression) else ())


@dataclass(frozen=True)
class IsDynamicSemistochastic(StochasticStyle):
    """Exact spawning from heavily populated configurations, stochastic otherwise, late
    compression (styles.jl:175-214)."""
    threshold: float = 1.0
    rel_spawning_threshold: float = 1.0
    abs_spawning_threshold: float = math.inf
    late_compression: bool = True
    kind = _native.STYLE_DYNAMIC

    @property
    def compression(self):
        return ThresholdCompression(self.threshold) if self.late_compression else NoCompression()

    @property
    def proj_threshold(self):
        return 0.0 if self.late_compression else float(self.threshold)

    def step_stat_names(self):
        return ("exact_steps", "inexact_steps", "spawn_attempts", "spawns") + self.compression.step_stat_names()

    def stat_values(self, st):
        base = (st.exact_steps, st.inexact_steps, st.spawn_attempts, st.spawns)
        return base + ((st.len_before,) if self.late_compression else ())


def default_style(value_type):
    """``default_style`` (styles.jl:216-219)."""
    return IsStochasticInteger() if value_type is int else IsDeterministic()
