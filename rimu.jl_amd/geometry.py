"""Lattice geometries of ``HubbardRealSpace``: ``CubicGrid`` and its constructors
``PeriodicBoundaries``, ``HardwallBoundaries``, ``LadderBoundaries``
(``src/Hamiltonians/geometry.jl:1-242``).

Sites are linear indices with the first dimension fastest (Julia's ``LinearIndices``);
``neighbor_site(site, chosen)`` moves along ``Directions(D)[chosen]`` (+e_1..+e_D, then
-e_1..-e_D) and folds periodic dimensions with ``mod1``; a move across a hard wall gives 0.
The device takes the whole neighbour table in ``rimu_model.nbr``.
"""
from __future__ import annotations

import math


class CubicGrid:
    """``CubicGrid(dims, fold)`` (geometry.jl:38-52): every dimension at least 2 long."""

    def __init__(self, dims, fold=None):
        dims = tuple(int(d) for d in (dims if isinstance(dims, (tuple, list)) else (dims,)))
        if any(d <= 1 for d in dims):
            raise ValueError("All dimensions must be at least 2 in size")
        fold = tuple(bool(f) for f in fold) if fold is not None else (True,) * len(dims)
        if len(fold) != len(dims):
            raise ValueError("fold needs one entry per dimension")
        self.dims = dims
        self.fold = fold

    def size(self):
        return self.dims

    def __len__(self):
        return math.prod(self.dims)

    def num_dimensions(self) -> int:
        return len(self.dims)

    def coords(self, site: int):
        """Cartesian coordinates (1-based) of a 1-based linear site index."""
        r, out = site - 1, []
        for n in self.dims:
            out.append(r % n + 1)
            r //= n
        return tuple(out)

    def index(self, coords) -> int:
        """Linear 1-based index of coordinates, folded where periodic; 0 if outside."""
        lin, mul = 0, 1
        for x, n, f in zip(coords, self.dims, self.fold):
            if f:
                x = (x - 1) % n + 1
            elif not 1 <= x <= n:
                return 0
            lin += (x - 1) * mul
            mul *= n
        return lin + 1

    def neighbor_site(self, site: int, chosen: int) -> int:
        """``neighbor_site(g, site, chosen)`` (geometry.jl:236-240)."""
        D = len(self.dims)
        x = list(self.coords(site))
        axis = (chosen - 1) % D
        x[axis] += 1 if chosen <= D else -1
        return self.index(x)

    def neighbour_table(self):
        """[site - 1][chosen - 1] -> neighbour site - 1, or 255 past a hard wall."""
        D = len(self.dims)
        return [[self.neighbor_site(s, c) - 1 if self.neighbor_site(s, c) else 255
                 for c in range(1, 2 * D + 1)] for s in range(1, len(self) + 1)]

    def __eq__(self, other):
        return isinstance(other, CubicGrid) and (self.dims, self.fold) == (other.dims, other.fold)

    def __repr__(self):
        return f"CubicGrid({self.dims}, {self.fold})"


def PeriodicBoundaries(*dims) -> CubicGrid:
    """All dimensions periodic (geometry.jl:58-66)."""
    dims = dims[0] if len(dims) == 1 and isinstance(dims[0], (tuple, list)) else dims
    return CubicGrid(dims, (True,) * len(dims))


def HardwallBoundaries(*dims) -> CubicGrid:
    """No dimension periodic (geometry.jl:68-77)."""
    dims = dims[0] if len(dims) == 1 and isinstance(dims[0], (tuple, list)) else dims
    return CubicGrid(dims, (False,) * len(dims))


def LadderBoundaries(*dims) -> CubicGrid:
    """The first dimension hard-walled, the others periodic: ``ntuple(>(1), Val(D))``
    (geometry.jl:79-88)."""
    dims = dims[0] if len(dims) == 1 and isinstance(dims[0], (tuple, list)) else dims
    return CubicGrid(dims, tuple(i > 0 for i in range(len(dims))))
