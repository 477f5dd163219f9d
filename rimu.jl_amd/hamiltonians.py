"""``HubbardReal1D``, ``HubbardMom1D`` and ``HubbardRealSpace`` host descriptors with the
Hamiltonian interface (``diagonal_element``, ``num_offdiagonals``, ``get_offdiagonal``,
``offdiagonals``, ``starting_address``, ``dimension``) evaluated on the device by
``rimu_hamiltonian_eval``.

References: ``src/Hamiltonians/HubbardReal1D.jl:22-62``, ``src/Hamiltonians/HubbardMom1D.jl:1-65``
(momentum grid and dispersion), ``src/Hamiltonians/HubbardRealSpace.jl:130-391``,
``src/Interfaces/hamiltonians.jl`` (interface).
"""
from __future__ import annotations

import math
from fractions import Fraction

import numpy as np

from . import _native
from .addresses import BoseFS, CompositeFS, FermiFS, key_words, pack_keys, unpack_keys
from .geometry import CubicGrid, PeriodicBoundaries


class AbstractHamiltonian:
    address: BoseFS
    u: float
    t: float

    def starting_address(self) -> BoseFS:
        return self.address

    @property
    def M(self) -> int:
        return self.address.M

    @property
    def N(self) -> int:
        return self.address.N

    def dimension(self) -> int:
        """``number_conserving_bose_dimension`` (Hamiltonians/abstract.jl:110)."""
        return math.comb(self.N + self.M - 1, self.N)

    def descriptor(self) -> _native.Model:
        raise NotImplementedError

    def _check(self, addr):
        if getattr(addr, "signature", None) != self.address.signature:
            raise ValueError("address does not match the Hamiltonian's address type, N and M")

    @property
    def key_words(self) -> int:
        return self.address.key_words

    def _keys(self, bits_list):
        return pack_keys(bits_list, self.key_words)

    # Interface -------------------------------------------------------------------------
    def diagonal_element(self, addr: BoseFS) -> float:
        self._check(addr)
        d, _, _, _ = _native.hamiltonian_eval(self.descriptor(), self._keys([addr.bits]), [0])
        return float(d[0])

    def num_offdiagonals(self, addr: BoseFS) -> int:
        self._check(addr)
        _, n, _, _ = _native.hamiltonian_eval(self.descriptor(), self._keys([addr.bits]), [0])
        return int(n[0])

    def get_offdiagonal(self, addr: BoseFS, chosen: int):
        self._check(addr)
        _, n, na, ov = _native.hamiltonian_eval(self.descriptor(), self._keys([addr.bits]), [chosen])
        if not 1 <= chosen <= n[0]:
            raise IndexError(f"chosen={chosen} outside 1:{n[0]}")
        return self.address.with_bits(unpack_keys(na)[0]), float(ov[0])

    def offdiagonals(self, addr: BoseFS):
        self._check(addr)
        n = self.num_offdiagonals(addr)
        if n == 0:
            return []
        _, _, na, ov = _native.hamiltonian_eval(self.descriptor(), self._keys([addr.bits] * n),
                                                list(range(1, n + 1)))
        return [(self.address.with_bits(a), float(v)) for a, v in zip(unpack_keys(na), ov)]

    def column(self, addrs_bits):
        """Diagonal elements and all off-diagonal (row, value) pairs of the columns of a batch
        of addresses (a word array, (n,) or (n, 2)): returns (diag, rows, cols, vals) with rows
        a word array of the same kind."""
        addrs_bits = np.asarray(addrs_bits, dtype=np.uint64)
        if self.key_words == 2:
            addrs_bits = addrs_bits.reshape(-1, 2)
        d, nod, _, _ = _native.hamiltonian_eval(self.descriptor(), addrs_bits,
                                                np.zeros(len(addrs_bits), dtype=np.int64))
        cols = np.repeat(np.arange(len(addrs_bits)), nod)
        chosen = np.concatenate([np.arange(1, n + 1) for n in nod]) if len(nod) else np.zeros(0, np.int64)
        if len(cols) == 0:
            return d, np.zeros((0, 2) if self.key_words == 2 else 0, np.uint64), cols, np.zeros(0)
        _, _, na, ov = _native.hamiltonian_eval(self.descriptor(), addrs_bits[cols], chosen)
        return d, na, cols, ov


class HubbardReal1D(AbstractHamiltonian):
    """``HubbardReal1D(address; u=1.0, t=1.0)``, periodic chain (HubbardReal1D.jl:22-29)."""

    def __init__(self, address: BoseFS, u: float = 1.0, t: float = 1.0):
        if address.M < 2:
            raise ValueError("HubbardReal1D needs at least two sites")
        self.address = address
        self.u = float(u)
        self.t = float(t)

    def descriptor(self) -> _native.Model:
        m = _native.Model()
        m.kind, m.M, m.N, m.u, m.t = _native.MODEL_REAL1D, self.M, self.N, self.u, self.t
        return m

    def __repr__(self):
        return f"HubbardReal1D({self.address!r}; u={self.u}, t={self.t})"

    def with_params(self, u=None, t=None) -> "HubbardReal1D":
        return HubbardReal1D(self.address, self.u if u is None else u, self.t if t is None else t)


def hubbard_dispersion(t: float, k: float) -> float:
    """``hubbard_dispersion`` (HubbardMom1D.jl:7): -2t cos(k)."""
    return (-2 * t) * math.cos(k)


class HubbardMom1D(AbstractHamiltonian):
    """``HubbardMom1D(address; u=1.0, t=1.0)`` (HubbardMom1D.jl:49-65)."""

    def __init__(self, address: BoseFS, u: float = 1.0, t: float = 1.0, dispersion=hubbard_dispersion):
        M = address.M
        if M < 3:
            raise ValueError("HubbardMom1D needs at least three modes")
        self.address = address
        self.u = float(u)
        self.t = float(t)
        step = 2 * math.pi / M
        start = (-math.pi * (1 + 1 / M) + step) if M % 2 == 1 else (-math.pi + step)
        # Julia's range(start; step, length=M) for Float64: correctly rounded start + j*step
        self.ks = np.array([float(Fraction(start) + j * Fraction(step)) for j in range(M)])
        self.kes = np.array([dispersion(self.t, k) for k in self.ks])

    def descriptor(self) -> _native.Model:
        m = _native.Model()
        m.kind, m.M, m.N, m.u, m.t = _native.MODEL_MOM1D, self.M, self.N, self.u, self.t
        for j, e in enumerate(self.kes):
            m.kes[j] = float(e)
        return m

    def __repr__(self):
        return f"HubbardMom1D({self.address!r}; u={self.u}, t={self.t})"


def _component_dimension(c) -> int:
    if isinstance(c, FermiFS):
        return math.comb(c.M, c.N)
    return math.comb(c.N + c.M - 1, c.N)


class HubbardRealSpace(AbstractHamiltonian):
    """``HubbardRealSpace(address; geometry, t, u, v)`` (HubbardRealSpace.jl:130-230) on a
    BoseFS, a FermiFS or a two-component CompositeFS.  t: one hopping per component; u: the
    symmetric C x C interaction matrix (a FermiFS component has no self-interaction); v:
    C x D trap strengths.  The default geometry is PeriodicBoundaries(M)."""

    def __init__(self, address, geometry: CubicGrid | None = None, t=None, u=None, v=None):
        comps = address.components if isinstance(address, CompositeFS) else (address,)
        C = len(comps)
        geometry = geometry if geometry is not None else PeriodicBoundaries(address.M)
        D = geometry.num_dimensions()
        t = [1.0] * C if t is None else [float(x) for x in np.ravel(t)]
        u = np.ones((C, C)) if u is None else np.array(u, dtype=float)
        v = np.zeros((C, D)) if v is None else np.array(v, dtype=float)
        if len(geometry) != address.M:
            raise ValueError("`geometry` does not have the correct number of sites")
        if u.size != C * C:
            raise ValueError(f"`u` must be a {C} × {C} matrix")
        u = u.reshape(C, C)
        if u.size != 1 and not np.array_equal(u, u.T):
            raise ValueError("`u` must be symmetric")
        if len(t) != C:
            raise ValueError(f"`t` must be a vector of length {C}")
        if v.size != C * D:
            raise ValueError(f"`v` must be a {C} × {D} matrix")
        if address.M > 64 or D > 3:
            raise ValueError("the device takes at most 64 sites in at most 3 dimensions")
        for c, comp in enumerate(comps):
            if isinstance(comp, FermiFS) and not np.all(u == 1) and u[c, c] != 0:
                import warnings
                warnings.warn(f"component {c + 1} is fermionic, but was given a self-interaction "
                              f"strength of {u[c, c]}")
        self.address = address
        self.geometry = geometry
        self.t = t
        self.u = u
        self.v = v.reshape(C, D)

    def dimension(self) -> int:
        """``number_conserving_dimension`` (Hamiltonians/abstract.jl:100-118)."""
        comps = self.address.components if isinstance(self.address, CompositeFS) else (self.address,)
        return math.prod(_component_dimension(c) for c in comps)

    def potential(self, c: int):
        """Trap potential of component c per site: sum_d v[c, d] x_d^2 with x_d in
        range(-fld(n_d, 2); length = n_d) (HubbardRealSpace.jl:214-224)."""
        out = []
        for s in range(1, self.M + 1):
            x = [xi - 1 - n // 2 for xi, n in zip(self.geometry.coords(s), self.geometry.dims)]
            acc = None
            for d, xd in enumerate(x):
                term = float(self.v[c, d]) * (xd * xd)
                acc = term if acc is None else acc + term
            out.append(acc)
        return out

    def descriptor(self) -> _native.Model:
        comps = self.address.components if isinstance(self.address, CompositeFS) else (self.address,)
        m = _native.Model()
        m.kind, m.M, m.N = _native.MODEL_REALSPACE, self.M, comps[0].N
        m.ncomp = len(comps)
        m.N2 = comps[1].N if len(comps) == 2 else 0
        m.fermi = sum(1 << c for c, x in enumerate(comps) if isinstance(x, FermiFS))
        m.ndir = 2 * self.geometry.num_dimensions()
        m.t, m.u = self.t[0], float(self.u[0, 0])
        if len(comps) == 2:
            m.t2, m.u12, m.u22 = self.t[1], float(self.u[0, 1]), float(self.u[1, 1])
        m.has_u = int(bool(np.any(self.u != 0)))
        m.has_pot = int(bool(np.any(self.v != 0)))
        if m.has_pot:
            for j, x in enumerate(self.potential(0)):
                m.kes[j] = x
            if len(comps) == 2:
                for j, x in enumerate(self.potential(1)):
                    m.pot2[j] = x
        for s, row in enumerate(self.geometry.neighbour_table()):
            for d, x in enumerate(row):
                m.nbr[s][d] = x
        return m

    def __repr__(self):
        return (f"HubbardRealSpace({self.address!r}; geometry={self.geometry}, t={self.t}, "
                f"u={self.u.tolist()}" + (f", v={self.v.tolist()}" if np.any(self.v != 0) else "") + ")")


def diagonal_element(h: AbstractHamiltonian, addr: BoseFS) -> float:
    return h.diagonal_element(addr)


def num_offdiagonals(h: AbstractHamiltonian, addr: BoseFS) -> int:
    return h.num_offdiagonals(addr)


def get_offdiagonal(h: AbstractHamiltonian, addr: BoseFS, chosen: int):
    return h.get_offdiagonal(addr, chosen)


def offdiagonals(h: AbstractHamiltonian, addr: BoseFS):
    return h.offdiagonals(addr)


def starting_address(h: AbstractHamiltonian) -> BoseFS:
    return h.starting_address()


def dimension(h: AbstractHamiltonian) -> int:
    return h.dimension()
