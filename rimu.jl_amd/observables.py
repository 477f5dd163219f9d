"""Operators for three-way dots and replica overlaps, and ``dot`` on device vectors.

* ``G2RealCorrelator(d)`` -- (1/M) sum_i n_i (n_{i+d} - delta_{0d}), periodic
  (``src/Hamiltonians/correlation_functions.jl:35-92``);
* ``DensityMatrixDiagonal(mode)`` -- n_mode (``src/Hamiltonians/DensityMatrixDiagonal.jl:1-44``);
* the Hamiltonians (``HubbardReal1D``, ``HubbardMom1D``) act as operators with off-diagonals.

``dot(t, u)``, ``dot(t, op, u)`` and ``dot(t, (op1, op2, ...), u)`` follow
``LinearAlgebra.dot`` on DVecs and ``dot_from_right`` (``abstractdvec.jl:313-324``,
``pdvec.jl:833-894``); on device vectors they run as one fused reduction
(``rimu_overlaps``).
"""
from __future__ import annotations

from . import _native
from .addresses import BoseFS
from .hamiltonians import AbstractHamiltonian


class AbstractObservable:
    def operator_descriptor(self, M: int, N: int) -> _native.Operator:
        raise NotImplementedError


class G2RealCorrelator(AbstractObservable):
    """Density-density correlation at distance d (diagonal)."""

    def __init__(self, d: int):
        self.d = int(d)

    def diagonal_element(self, addr: BoseFS) -> float:
        onr = addr.onr
        M = len(onr)
        dd = self.d % M
        acc = sum(onr[i] * (onr[(i + dd) % M] - (dd == 0)) for i in range(M))
        return acc / M

    def operator_descriptor(self, M, N):
        if not -M < self.d < M:
            raise ValueError("G2RealCorrelator needs |d| < M")
        op = _native.Operator()
        op.kind, op.param = _native.OP_G2_REAL, self.d
        return op

    def __eq__(self, other):
        return isinstance(other, G2RealCorrelator) and other.d == self.d

    def __hash__(self):
        return hash(("G2RealCorrelator", self.d))

    def __repr__(self):
        return f"G2RealCorrelator({self.d})"


class DensityMatrixDiagonal(AbstractObservable):
    """Occupation number of one mode (1-based), a diagonal operator."""

    def __init__(self, mode: int, component: int = 0):
        if component not in (0, 1):
            raise ValueError("single-component addresses have component 0 or 1")
        self.mode = int(mode)
        self.component = component

    def diagonal_element(self, addr: BoseFS) -> float:
        return float(addr.onr[self.mode - 1])

    def operator_descriptor(self, M, N):
        if not 1 <= self.mode <= M:
            raise ValueError("DensityMatrixDiagonal mode outside 1..M")
        op = _native.Operator()
        op.kind, op.param = _native.OP_DENSITY_DIAGONAL, self.mode
        return op

    def __eq__(self, other):
        return isinstance(other, DensityMatrixDiagonal) and other.mode == self.mode

    def __hash__(self):
        return hash(("DensityMatrixDiagonal", self.mode))

    def __repr__(self):
        return f"DensityMatrixDiagonal({self.mode})"


def operator_descriptor(op, M: int, N: int) -> _native.Operator:
    if isinstance(op, AbstractHamiltonian):
        if op.M != M or op.N != N:
            raise ValueError("operator acts on another address space")
        d = _native.Operator()
        d.kind, d.model = _native.OP_HAMILTONIAN, op.descriptor()
        return d
    if isinstance(op, AbstractObservable):
        return op.operator_descriptor(M, N)
    raise TypeError(f"{op!r} is not an operator the device supports")


def _engine_of(v):
    if hasattr(v, "state_vector"):
        v = v.state_vector
    if hasattr(v, "engine"):
        return v.engine, v.hamiltonian.M, v.hamiltonian.N
    raise TypeError("dot on the device takes DeviceDVecs (or simulations)")


def overlap_values(t, u, ops=()) -> list:
    """[dot(t, u), dot(t, op, u) for op in ops], in batches of at most 8 operators."""
    te, M, N = _engine_of(t)
    ue, M2, N2 = _engine_of(u)
    if (M, N) != (M2, N2):
        raise ValueError("vectors of different address spaces")
    descs = [operator_descriptor(op, M, N) for op in ops]
    out = []
    first = True
    for i in range(0, max(len(descs), 1), _native.MAX_OPERATORS):
        chunk = descs[i:i + _native.MAX_OPERATORS]
        vals = _native.overlaps(te, ue, chunk)
        if first:
            out.append(float(vals[0]))
            first = False
        out.extend(float(x) for x in vals[1:])
    return out


def dot(t, *args):
    """``dot(t, u)``, ``dot(t, op, u)`` or ``dot(t, (ops...), u)`` of device vectors."""
    if len(args) == 1:
        return overlap_values(t, args[0])[0]
    if len(args) != 2:
        raise TypeError("dot(t, u) or dot(t, op, u)")
    op, u = args
    if isinstance(op, (tuple, list)):
        return tuple(overlap_values(t, u, op)[1:])
    return overlap_values(t, u, (op,))[1]
