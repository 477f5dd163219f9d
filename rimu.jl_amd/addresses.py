"""``BoseFS`` addresses on one 64-bit chunk.

Layout of ``BitStringAddresses/bitstring.jl:464-471`` (``from_bose_onr``): mode 1 in the
least significant bits, mode i written as n_i one-bits followed by a zero separator, B =
N + M - 1 bits in total.  Constructors follow ``BitStringAddresses/bosefs.jl:54-112``.
"""
from __future__ import annotations

from dataclasses import dataclass


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
        if len(onr) + sum(onr) - 1 > 64:
            raise ValueError("address does not fit one 64-bit chunk (N + M - 1 > 64)")
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

    def num_occupied_modes(self) -> int:
        return sum(1 for n in self.onr if n > 0)

    def __repr__(self):
        return "fs\"|" + " ".join(str(n) for n in self.onr) + "⟩\""

    def __lt__(self, other):  # isless is numeric on the chunk (bitstring.jl:317-324)
        return self.bits < other.bits


def near_uniform(N: int, M: int) -> BoseFS:
    """``near_uniform(BoseFS{N,M})`` (bosefs.jl:139-160)."""
    fill, extras = divmod(N, M)
    onr = [fill] * M
    for i in range(extras):
        onr[i] += 1
    return BoseFS(tuple(onr))
