"""Vectors: ``DVec`` (host dictionary, for building start vectors and reading results) and
``DeviceDVec`` (the ``AbstractDVec`` whose storage is the engine's sorted array in HBM).

References: ``src/DictVectors/dvec.jl:44-136`` (DVec, zero deletes the entry),
``src/Interfaces/dictvectors.jl:40-51`` (``AbstractDVec``, ``deposit!``),
``src/DictVectors/abstractdvec.jl:258-270`` (``walkernumber``, ``walkernumber_and_length``),
``src/DictVectors/projectors.jl:140-173`` (``FrozenDVec`` and its ``dot``),
``src/DictVectors/pdvec.jl:156-200`` and ``src/DictVectors/initiatordvec.jl:42-110``
(the ``initiator`` keyword of ``PDVec`` and ``InitiatorDVec``).
"""
from __future__ import annotations

import numpy as np

from .addresses import BoseFS, CompositeFS, FermiFS, key_words, pack_keys, unpack_keys

ADDRESS_TYPES = (BoseFS, FermiFS, CompositeFS)
from .initiators import Initiator, InitiatorRule, NonInitiator, initiator_rule
from .styles import IsStochasticInteger, StochasticStyle, default_style


class DVec:
    """Dictionary-backed sparse vector (dvec.jl:44).  Setting a value to zero deletes it.
    ``initiator`` is the vector's initiator rule (NonInitiator for a DVec)."""

    initiator: InitiatorRule = NonInitiator()

    def __init__(self, *pairs, style: StochasticStyle | None = None, capacity: int = 0):
        self.storage = {}
        items = []
        for p in pairs:
            if isinstance(p, dict):
                items.extend(p.items())
            elif isinstance(p, DVec):
                items.extend(p.storage.items())
            elif isinstance(p, (list, tuple)) and len(p) and isinstance(p[0], (list, tuple)):
                items.extend(p)
            else:
                items.append(p)
        if style is None:
            vt = int if all(isinstance(v, (int, np.integer)) for _, v in items) and items else float
            style = default_style(vt)
        self.style = style
        for k, v in items:
            self[k] = v

    def _conv(self, v):
        return int(v) if isinstance(self.style, IsStochasticInteger) else float(v)

    def __getitem__(self, k):
        return self.storage.get(k, self._conv(0))

    def __setitem__(self, k, v):
        v = self._conv(v)
        if v == 0:
            self.storage.pop(k, None)
        else:
            self.storage[k] = v

    def deposit(self, k, v):
        """``deposit!`` (Interfaces/dictvectors.jl:49-51)."""
        self[k] = self[k] + v

    def __len__(self):
        return len(self.storage)

    def pairs(self):
        return self.storage.items()

    def keys(self):
        return self.storage.keys()

    def values(self):
        return self.storage.values()

    def walkernumber(self) -> float:
        return float(sum(abs(v) for v in self.storage.values()))

    def copy(self):
        out = DVec(dict(self.storage), style=self.style)
        out.initiator = self.initiator
        return out

    def to_arrays(self):
        """Sorted word array of the keys ((n,) or (n, 2)) and the values."""
        ks = sorted(self.storage)
        kw = ks[0].key_words if ks else 1
        return pack_keys([k.bits for k in ks], kw), np.array([self.storage[k] for k in ks])

    def __repr__(self):
        return f"{type(self).__name__} with {len(self)} entries, style = {self.style}, initiator = {self.initiator}"


class PDVec(DVec):
    """``PDVec(pairs...; style, initiator, initiator_threshold)`` (pdvec.jl:156-200): the
    host container of a start vector with an initiator rule; ``initiator`` is ``False``
    (default), ``True`` (``Initiator(initiator_threshold)``, threshold 1 when 0) or a rule."""

    def __init__(self, *pairs, style: StochasticStyle | None = None, initiator=None,
                 initiator_threshold=0, capacity: int = 0):
        super().__init__(*pairs, style=style, capacity=capacity)
        self.initiator = initiator_rule(initiator, initiator_threshold)


class InitiatorDVec(DVec):
    """``InitiatorDVec(pairs...; style, initiator, initiator_threshold)``
    (initiatordvec.jl:42-110): default rule ``Initiator(initiator_threshold)`` with
    threshold 1."""

    def __init__(self, *pairs, style: StochasticStyle | None = None, initiator=None,
                 initiator_threshold=1, capacity: int = 0):
        super().__init__(*pairs, style=style, capacity=capacity)
        self.initiator = Initiator(initiator_threshold) if initiator is None else initiator


class FrozenDVec:
    """``freeze(dv)`` (projectors.jl:140-173): a list of pairs with a fast ``dot``."""

    def __init__(self, pairs):
        self.pairs = list(pairs)

    def keys_bits(self):
        kw = self.pairs[0][0].key_words if self.pairs else 1
        return pack_keys([k.bits for k, _ in self.pairs], kw)

    def coefs(self):
        return np.array([float(v) for _, v in self.pairs], dtype=np.float64)


def freeze(v) -> FrozenDVec:
    if isinstance(v, FrozenDVec):
        return v
    if isinstance(v, ADDRESS_TYPES):
        return FrozenDVec([(v, 1.0)])
    return FrozenDVec(list(v.pairs()))


class DeviceDVec:
    """The state vector of a running simulation: sorted keys and values held by a
    ``_native.Engine`` in device memory.  Reading it (``pairs``, ``getindex``) copies to
    the host."""

    def __init__(self, engine, hamiltonian, style, initiator: InitiatorRule = NonInitiator()):
        self.engine = engine
        self.hamiltonian = hamiltonian
        self.style = style
        self.initiator = initiator
        self._norm = None

    def __len__(self):
        return len(self.engine)

    def walkernumber(self) -> float:
        if self._norm is None:
            _, v = self.engine.download()
            self._norm = float(np.abs(v).sum())
        return self._norm

    def walkernumber_and_length(self):
        return self.walkernumber(), len(self)

    def arrays(self):
        return self.engine.download()

    def to_dvec(self) -> DVec:
        keys, vals = self.engine.download()
        template = self.hamiltonian.address
        out = DVec(style=self.style)
        for k, v in zip(unpack_keys(keys), vals.tolist()):
            out.storage[template.with_bits(k)] = v
        return out

    def pairs(self):
        return self.to_dvec().pairs()

    def __getitem__(self, addr: BoseFS):
        keys, vals = self.engine.download()
        index = {k: i for i, k in enumerate(unpack_keys(keys))}
        i = index.get(addr.bits)
        return vals[i].item() if i is not None else 0

    def __repr__(self):
        return f"DeviceDVec with {len(self)} entries, style = {self.style}, initiator = {self.initiator}"
