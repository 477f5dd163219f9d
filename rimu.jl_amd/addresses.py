"""Fock addresses on one or two 64-bit chunks: ``BoseFS``, ``FermiFS`` and two-component
``CompositeFS`` (``FermiFS2C``).

BoseFS layout of ``BitStringAddresses/bitstring.jl:464-508`` (``from_bose_onr``): mode 1 in
the least significant bits, mode i written as n_i one-bits followed by a zero separator,
B = N + M - 1 bits.  FermiFS: one bit per mode, mode j at bit j-1
(``from_fermi_onr``, bitstring.jl:713-723), B = M.  CompositeFS: component 1's bits above
component 2's, so numeric order is the tuple order of ``isless`` (multicomponent.jl:79).
At most 128 bits in all.  Constructors follow ``bosefs.jl:54-112``, ``fermifs.jl:1-100``
and ``multicomponent.jl:1-148``.  Arrays of addresses for the device are uint64 words:
shape (n,) for B <= 64, (n, 2) with the low word first for 64 < B <= 128.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

MASK64 = (1 << 64) - 1


def key_words(M: int, N: int) -> int:
    """64-bit words per address of N bosons in M modes: 1 or 2."""
    return 2 if M + N - 1 > 64 else 1


def words_for_bits(B: int) -> int:
    """64-bit words per address of B bits: 1 or 2."""
    return 2 if B > 64 else 1


def pack_keys(ints, kw: int) -> np.ndarray:
    """Python-int bit strings as a word array ((n,) or (n, 2) [lo, hi])."""
    ints = [int(x) for x in ints]
    if kw == 1:
        return np.array(ints, dtype=np.uint64)
    return np.array([[x & MASK64, x >> 64] for x in ints], dtype=np.uint64).reshape(len(ints), 2)


def unpack_keys(arr) -> list:
    """The inverse of pack_keys: a list of Python ints."""
    arr = np.asarray(arr, dtype=np.uint64)
    if arr.ndim == 1:
        return [int(x) for x in arr.tolist()]
    return [int(lo) | (int(hi) << 64) for lo, hi in arr.tolist()]


def sort_keys(keys: np.ndarray) -> np.ndarray:
    """Stable ascending order of a word array (numeric order of the bit strings)."""
    if keys.ndim == 1:
        return np.argsort(keys, kind="stable")
    return np.lexsort((keys[:, 0], keys[:, 1]))


def onr_to_bits(onr) -> int:
    """``from_bose_onr`` (bitstring.jl:464-471)."""
    r = 0
    for n in reversed(list(onr)):
        n = int(n)
        if n < 0:
            raise ValueError("negative occupation number")
        r = (r << (n + 1)) | ((1 << n) - 1)
    return r


def bits_to_onr(bits: int, M: int):
    """``to_bose_onr`` (bitstring.jl:511-520)."""
    out = []
    for _ in range(M):
        n = 0
        while bits & 1:
            n += 1
            bits >>= 1
        out.append(n)
        bits >>= 1
    return tuple(out)


@dataclass(frozen=True)
class BoseFS:
    """Fock state of N bosons in M modes (``BoseFS{N,M}``, bosefs.jl:54)."""
    N: int
    M: int
    bits: int

    def __init__(self, *args, M=None, N=None, bits=None):
        if bits is not None:
            if M is None or N is None:
                raise ValueError("BoseFS(bits=...) needs M and N")
            object.__setattr__(self, "N", int(N))
            object.__setattr__(self, "M", int(M))
            object.__setattr__(self, "bits", int(bits))
            return
        if len(args) == 1 and isinstance(args[0], (tuple, list)):
            onr = tuple(int(x) for x in args[0])
        else:
            onr = tuple(int(x) for x in args)
        if M is not None and len(onr) != M:
            raise ValueError(f"invalid ONR: {M} modes expected, {len(onr)} given")
        if N is not None and sum(onr) != N:
            raise ValueError(f"invalid ONR: {N} particles expected, {sum(onr)} given")
        if len(onr) + sum(onr) - 1 > 128:
            raise ValueError("address does not fit two 64-bit chunks (N + M - 1 > 128)")
        object.__setattr__(self, "N", sum(onr))
        object.__setattr__(self, "M", len(onr))
        object.__setattr__(self, "bits", onr_to_bits(onr))

    @classmethod
    def from_pairs(cls, M: int, pairs) -> "BoseFS":
        """``BoseFS(M, mode => n, ...)`` (bosefs.jl:109-112)."""
        onr = [0] * M
        items = pairs.items() if hasattr(pairs, "items") else pairs
        for mode, n in items:
            onr[int(mode) - 1] += int(n)
        return cls(tuple(onr))

    @property
    def onr(self):
        return bits_to_onr(self.bits, self.M)

    @property
    def key_bits(self) -> int:
        return self.N + self.M - 1

    @property
    def key_words(self) -> int:
        return key_words(self.M, self.N)

    @property
    def signature(self):
        return ("BoseFS", self.N, self.M)

    def with_bits(self, bits: int) -> "BoseFS":
        """The address of the same type with bit string ``bits``."""
        return BoseFS(M=self.M, N=self.N, bits=bits)

    def num_occupied_modes(self) -> int:
        return sum(1 for n in self.onr if n > 0)

    def __repr__(self):
        return "fs\"|" + " ".join(str(n) for n in self.onr) + "⟩\""

    def __lt__(self, other):  # isless is numeric on the chunk (bitstring.jl:317-324)
        return self.bits < other.bits


@dataclass(frozen=True)
class FermiFS:
    """Fock state of N fermions in M modes (``FermiFS{N,M}``, fermifs.jl:1-100)."""
    N: int
    M: int
    bits: int

    def __init__(self, *args, M=None, N=None, bits=None):
        if bits is not None:
            if M is None or N is None:
                raise ValueError("FermiFS(bits=...) needs M and N")
            object.__setattr__(self, "N", int(N))
            object.__setattr__(self, "M", int(M))
            object.__setattr__(self, "bits", int(bits))
            return
        onr = tuple(int(x) for x in (args[0] if len(args) == 1 and isinstance(args[0], (tuple, list)) else args))
        if any(n not in (0, 1) for n in onr):
            raise ValueError("Invalid ONR: may only contain 0s and 1s.")
        if M is not None and len(onr) != M:
            raise ValueError(f"Invalid ONR: {M} modes expected, {len(onr)} given.")
        if N is not None and sum(onr) != N:
            raise ValueError(f"Invalid ONR: {N} particles expected, {sum(onr)} given.")
        if len(onr) > 128:
            raise ValueError("address does not fit two 64-bit chunks (M > 128)")
        object.__setattr__(self, "N", sum(onr))
        object.__setattr__(self, "M", len(onr))
        object.__setattr__(self, "bits", sum(1 << i for i, n in enumerate(onr) if n))

    @classmethod
    def from_pairs(cls, M: int, pairs) -> "FermiFS":
        """``FermiFS(M, mode => 1, ...)`` (fermifs.jl:98-100)."""
        onr = [0] * M
        items = pairs.items() if hasattr(pairs, "items") else pairs
        for mode, n in items:
            onr[int(mode) - 1] += int(n)
        return cls(tuple(onr))

    @property
    def onr(self):
        return tuple((self.bits >> i) & 1 for i in range(self.M))

    @property
    def key_bits(self) -> int:
        return self.M

    @property
    def key_words(self) -> int:
        return words_for_bits(self.M)

    @property
    def signature(self):
        return ("FermiFS", self.N, self.M)

    def with_bits(self, bits: int) -> "FermiFS":
        return FermiFS(M=self.M, N=self.N, bits=bits)

    def num_occupied_modes(self) -> int:
        return self.N

    def __repr__(self):
        return "fs\"|" + "".join("↑" if n else "⋅" for n in self.onr) + "⟩\""

    def __lt__(self, other):
        return self.bits < other.bits


@dataclass(frozen=True)
class CompositeFS:
    """``CompositeFS(c1, c2)`` (multicomponent.jl:1-40): two single-component addresses with
    the same number of modes.  ``bits`` is component 1 above component 2."""
    components: tuple

    def __init__(self, *components):
        if len(components) == 1 and isinstance(components[0], tuple):
            components = components[0]
        if len(components) != 2:
            raise ValueError("the device takes CompositeFS addresses of two components")
        if len({c.M for c in components}) != 1:
            raise ValueError("all addresses must have the same number of modes")
        if sum(c.key_bits for c in components) > 128:
            raise ValueError("address does not fit two 64-bit chunks (more than 128 bits)")
        object.__setattr__(self, "components", tuple(components))

    @property
    def M(self) -> int:
        return self.components[0].M

    @property
    def N(self) -> int:
        return sum(c.N for c in self.components)

    @property
    def bits(self) -> int:
        a, b = self.components
        return (a.bits << b.key_bits) | b.bits

    @property
    def key_bits(self) -> int:
        return sum(c.key_bits for c in self.components)

    @property
    def key_words(self) -> int:
        return words_for_bits(self.key_bits)

    @property
    def signature(self):
        return ("CompositeFS",) + tuple(c.signature for c in self.components)

    @property
    def onr(self):
        return tuple(c.onr for c in self.components)

    def with_bits(self, bits: int) -> "CompositeFS":
        a, b = self.components
        w = b.key_bits
        return CompositeFS(a.with_bits(bits >> w), b.with_bits(bits & ((1 << w) - 1)))

    def num_occupied_modes(self) -> int:
        return sum(c.num_occupied_modes() for c in self.components)

    def __repr__(self):
        return "CompositeFS(" + ", ".join(repr(c) for c in self.components) + ")"

    def __lt__(self, other):  # isless on the tuple of components (multicomponent.jl:79)
        return self.bits < other.bits


def FermiFS2C(*args):
    """``FermiFS2C(onr_a, onr_b)``, ``FermiFS2C(f1, f2)`` or ``FermiFS2C(M, mode => +-1, ...)``
    (multicomponent.jl:89-140): +1 puts a fermion in component 1, -1 in component 2."""
    if len(args) == 2 and all(isinstance(a, FermiFS) for a in args):
        return CompositeFS(args[0], args[1])
    if len(args) == 2 and all(isinstance(a, (tuple, list)) for a in args):
        return CompositeFS(FermiFS(tuple(args[0])), FermiFS(tuple(args[1])))
    M, pairs = args[0], args[1:]
    if len(pairs) == 1 and not isinstance(pairs[0], tuple):
        pairs = tuple(pairs[0])
    up = [0] * M
    down = [0] * M
    for mode, n in pairs:
        if n > 0:
            up[mode - 1] += n
        elif n < 0:
            down[mode - 1] -= n
    return CompositeFS(FermiFS(tuple(up)), FermiFS(tuple(down)))


def near_uniform(N: int, M: int, kind=None):
    """``near_uniform(BoseFS{N,M})`` (bosefs.jl:139-160) or, with kind=FermiFS,
    ``near_uniform(FermiFS{N,M})`` (fermifs.jl:131-133: the first N modes filled)."""
    if kind is FermiFS:
        return FermiFS(tuple([1] * N + [0] * (M - N)))
    fill, extras = divmod(N, M)
    onr = [fill] * M
    for i in range(extras):
        onr[i] += 1
    return BoseFS(tuple(onr))
