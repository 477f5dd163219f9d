"""``BoseFS`` addresses on one or two 64-bit chunks.

Layout of ``BitStringAddresses/bitstring.jl:464-508`` (``from_bose_onr``): mode 1 in the
least significant bits, mode i written as n_i one-bits followed by a zero separator, B =
N + M - 1 bits in total (at most 128 here).  Constructors follow
``BitStringAddresses/bosefs.jl:54-112``.  Arrays of addresses for the device are uint64
words: shape (n,) for B <= 64, (n, 2) with the low word first for 64 < B <= 128.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

MASK64 = (1 << 64) - 1


def key_words(M: int, N: int) -> int:
    """64-bit words per address of N bosons in M modes: 1 or 2."""
    return 2 if M + N - 1 > 64 else 1


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
    def key_words(self) -> int:
        return key_words(self.M, self.N)

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
