"""Initiator rules: ``NonInitiator``, ``Initiator``, ``SimpleInitiator`` and
``CoherentInitiator`` (``src/DictVectors/initiators.jl:132-236``).

A rule travels with a vector (``PDVec(...; initiator=...)``, ``InitiatorDVec``) and is
applied on the device by the step engine (``rimu_engine_set_initiator``): every deposit is
classified by its parent as safe, unsafe or initiator (``to_initiator_value``,
``initiators.jl:143-157``) and every address keeps the rule's ``from_initiator_value`` of
its three components.  ``InitiatorValue`` and the host-side ``from_initiator_value`` below
state the same rules for reading and testing.
"""
from __future__ import annotations

from dataclasses import dataclass

from . import _native


@dataclass(frozen=True)
class InitiatorValue:
    """Composite walker (safe, unsafe, initiator), ``initiators.jl:22-45``."""
    safe: float = 0
    unsafe: float = 0
    initiator: float = 0

    def __add__(self, o: "InitiatorValue") -> "InitiatorValue":
        return InitiatorValue(self.safe + o.safe, self.unsafe + o.unsafe,
                              self.initiator + o.initiator)

    def iszero(self) -> bool:
        return self.safe == 0 and self.unsafe == 0 and self.initiator == 0


class InitiatorRule:
    code = _native.INITIATOR_NONE
    threshold = 0.0

    def descriptor(self):
        return self.code, float(self.threshold)

    def to_initiator_value(self, add, val, parent) -> InitiatorValue:
        """``_default_to_initiator_value``, initiators.jl:143-157."""
        p_add, p_val = parent
        big = abs(p_val) > self.threshold
        if p_add == add:
            return InitiatorValue(initiator=val) if big else InitiatorValue(safe=val)
        return InitiatorValue(safe=val) if big else InitiatorValue(unsafe=val)

    def __eq__(self, other):
        return type(self) is type(other) and self.threshold == other.threshold

    def __hash__(self):
        return hash((type(self).__name__, self.threshold))

    def __repr__(self):
        return f"{type(self).__name__}({self.threshold})"


class NonInitiator(InitiatorRule):
    """The approximation switched off (the default of ``PDVec``), initiators.jl:218-236."""

    def to_initiator_value(self, add, val, parent) -> InitiatorValue:
        return InitiatorValue(safe=val)

    def from_initiator_value(self, v: InitiatorValue):
        return v.safe

    def __repr__(self):
        return "NonInitiator()"


class Initiator(InitiatorRule):
    """Initiators (|v| > threshold) spawn anywhere; non-initiators spawn onto initiators
    only (initiators.jl:119-161)."""
    code = _native.INITIATOR_DEFAULT

    def __init__(self, threshold=1.0):
        self.threshold = threshold

    def from_initiator_value(self, v: InitiatorValue):
        return (v.safe + v.initiator) + (v.unsafe if v.initiator != 0 else 0)


class SimpleInitiator(InitiatorRule):
    """Non-initiators do not spawn at all (initiators.jl:163-186)."""
    code = _native.INITIATOR_SIMPLE

    def __init__(self, threshold=1.0):
        self.threshold = threshold

    def from_initiator_value(self, v: InitiatorValue):
        return v.safe + v.initiator


class CoherentInitiator(InitiatorRule):
    """As ``Initiator``, and unsafe spawns onto a non-initiator also survive when they add
    up to more than the threshold (initiators.jl:188-216)."""
    code = _native.INITIATOR_COHERENT

    def __init__(self, threshold=1.0):
        self.threshold = threshold

    def from_initiator_value(self, v: InitiatorValue):
        if v.initiator != 0 or abs(v.unsafe) > self.threshold:
            return v.initiator + v.safe + v.unsafe
        return v.initiator + v.safe


def initiator_rule(initiator=None, initiator_threshold=0) -> InitiatorRule:
    """The ``initiator`` / ``initiator_threshold`` keywords of ``PDVec`` (pdvec.jl:168-185):
    ``False`` -> NonInitiator, ``True`` -> Initiator(threshold, 1 when 0), a rule as given;
    by default ``initiator = initiator_threshold > 0``."""
    if initiator is None:
        initiator = initiator_threshold > 0
    if initiator is False:
        return NonInitiator()
    if initiator is True:
        return Initiator(initiator_threshold if initiator_threshold != 0 else 1)
    if isinstance(initiator, InitiatorRule):
        return initiator
    raise ValueError(f"Invalid initiator {initiator!r}")
