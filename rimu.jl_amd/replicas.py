"""Replica strategies: ``NoStats`` and ``AllOverlaps``
(``src/strategies_and_params/replicastrategy.jl:1-183``).

``AllOverlaps(n; operator, vecnorm)`` runs n replicas and reports, for every pair i < j in
that order, ``c{i}_dot_c{j}`` (unless ``vecnorm=False``) and ``c{i}_Op{k}_c{j}`` for each
operator k -- the overlaps ``dot(v_i, v_j)`` and ``dot(v_i, op_k, v_j)`` computed on the device
(``all_overlaps``, ``replicastrategy.jl:160-183``).
"""
from __future__ import annotations

from .hamiltonians import AbstractHamiltonian
from .observables import AbstractObservable, overlap_values


class ReplicaStrategy:
    n_replicas = 1

    def replica_stats(self, vectors):
        return (), ()


class NoStats(ReplicaStrategy):
    """n replicas, no replica statistics (replicastrategy.jl:50-63)."""

    def __init__(self, n_replicas: int = 1):
        self.n_replicas = int(n_replicas)

    def __eq__(self, other):
        return isinstance(other, NoStats) and other.n_replicas == self.n_replicas

    def __repr__(self):
        return f"NoStats({self.n_replicas})"


class AllOverlaps(ReplicaStrategy):
    """Overlaps of all replica pairs, with operators (replicastrategy.jl:65-145)."""

    def __new__(cls, n_replicas=2, operator=None, transform=None, vecnorm=True):
        if not isinstance(n_replicas, int):
            raise TypeError("n_replicas must be an integer")
        if transform is not None:
            raise NotImplementedError("transformed Hamiltonians are not supported")
        if operator is None:
            ops = ()
        elif isinstance(operator, (tuple, list)):
            if not all(isinstance(o, (AbstractHamiltonian, AbstractObservable)) for o in operator):
                raise TypeError("operator must be an operator or a tuple/list of operators")
            ops = tuple(operator)
        else:
            ops = (operator,)
        if not vecnorm and not ops:
            return NoStats(n_replicas)       # replicastrategy.jl:130-132
        self = super().__new__(cls)
        self.n_replicas = n_replicas
        self.operators = ops
        self.vecnorm = bool(vecnorm)
        return self

    def __init__(self, n_replicas=2, operator=None, transform=None, vecnorm=True):
        pass

    def replica_stats(self, vectors):
        names, values = [], []
        n = len(vectors)
        for i in range(n):
            for j in range(i + 1, n):
                vals = overlap_values(vectors[i], vectors[j], self.operators)
                if self.vecnorm:
                    names.append(f"c{i + 1}_dot_c{j + 1}")
                    values.append(vals[0])
                for k in range(len(self.operators)):
                    names.append(f"c{i + 1}_Op{k + 1}_c{j + 1}")
                    values.append(vals[1 + k])
        return tuple(names), tuple(values)

    def __repr__(self):
        return f"AllOverlaps({self.n_replicas}; operator={self.operators}, vecnorm={self.vecnorm})"
