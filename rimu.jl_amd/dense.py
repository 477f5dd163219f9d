"""Deterministic products on the ranked (full) basis and their adjoint.

``DenseHamiltonians`` holds the sorted basis of N bosons in M modes on the device together
with several ``HubbardReal1D`` Hamiltonians sharing M and N (one right-hand side each).
``apply`` is ``mul!(w, H, v)`` / one IsDeterministic step on full vectors
(``src/DictVectors/abstractdvec.jl:275-284``, ``src/StochasticStyles/styles.jl:101-105``);
``rrule_mul`` is the ChainRulesCore-style rule of that product (value plus pullback giving
the cotangents of v, u and t); ``projector_energy`` runs K projector steps and returns the
Rayleigh quotient (``src/Hamiltonians/abstract.jl:153``) with dE/du and dE/dt from the
checkpointed adjoint of the whole projection.  Vectors are numpy arrays of shape (dim, R).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native
from .hamiltonians import HubbardReal1D

_p = _native._p


def basis_size(M: int, N: int) -> int:
    d = ctypes.c_int64(0)
    _native.check(_native.lib().rimu_basis_size(M, N, ctypes.byref(d)))
    return int(d.value)


class DenseHamiltonians:
    def __init__(self, hamiltonians):
        hs = list(hamiltonians)
        if not hs or not all(isinstance(h, HubbardReal1D) for h in hs):
            raise TypeError("DenseHamiltonians takes HubbardReal1D Hamiltonians")
        self.hamiltonians = hs
        self.R = len(hs)
        arr = (_native.Model * self.R)(*[h.descriptor() for h in hs])
        self._h = ctypes.c_void_p()
        _native.check(_native.lib().rimu_dense_create(arr, self.R, ctypes.byref(self._h)))
        d = ctypes.c_int64(0)
        _native.check(_native.lib().rimu_dense_dim(self._h, ctypes.byref(d)))
        self.dim = int(d.value)

    def close(self):
        if self._h:
            _native.lib().rimu_dense_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def basis(self) -> np.ndarray:
        out = np.zeros(self.dim, dtype=np.uint64)
        _native.check(_native.lib().rimu_dense_basis(self._h, _p(out)))
        return out

    def _vec(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64).reshape(self.dim, self.R)
        return v

    def _coef(self, x):
        return np.ascontiguousarray(np.broadcast_to(np.asarray(x, dtype=np.float64), (self.R,)))

    def apply(self, a, b, v) -> np.ndarray:
        """w_i = (a_i + b_i H_i) v_i."""
        v = self._vec(v)
        w = np.empty_like(v)
        a, b = self._coef(a), self._coef(b)
        _native.check(_native.lib().rimu_dense_apply(self._h, _p(a), _p(b), _p(v), _p(w)))
        return w

    def pullback(self, a, b, v, ybar):
        """Cotangents (vbar, ubar, tbar) of y = (a + b H) v for the cotangent ybar."""
        v, ybar = self._vec(v), self._vec(ybar)
        vbar = np.empty_like(v)
        ubar = np.zeros(self.R)
        tbar = np.zeros(self.R)
        a, b = self._coef(a), self._coef(b)
        _native.check(_native.lib().rimu_dense_pullback(self._h, _p(a), _p(b), _p(v), _p(ybar),
                                                        _p(vbar), _p(ubar), _p(tbar)))
        return vbar, ubar, tbar

    def rrule_mul(self, v):
        """(H v, pullback) -- the rule of mul!(w, H, v) with respect to v, u and t."""
        y = self.apply(0.0, 1.0, v)

        def pullback(ybar):
            return self.pullback(0.0, 1.0, v, ybar)

        return y, pullback

    def projector_energy(self, K: int, dtau: float, shift, v0, ckpt: int | None = None):
        """K steps of v <- (1 + dtau (S - H)) v, then E = <v|H|v>/<v|v> and dE/du, dE/dt by
        the adjoint with recomputation from checkpoints every ``ckpt`` steps (default
        ceil(sqrt(K))).  Returns a dict of arrays over the right-hand sides."""
        if ckpt is None:
            ckpt = max(1, int(np.ceil(np.sqrt(max(K, 1)))))
        v0 = self._vec(v0)
        shift = self._coef(shift)
        outs = {k: np.zeros(self.R) for k in ("energy", "grad_u", "grad_t", "hf_u", "hf_t")}
        vf = np.empty_like(v0)
        _native.check(_native.lib().rimu_dense_projector(
            self._h, int(K), float(dtau), _p(shift), int(ckpt), _p(v0), _p(outs["energy"]),
            _p(outs["grad_u"]), _p(outs["grad_t"]), _p(outs["hf_u"]), _p(outs["hf_t"]), _p(vf)))
        outs["v"] = vf
        outs["ckpt"] = ckpt
        return outs

    def ground_state(self, v0=None, max_iters: int = 3000, tol: float = 1e-12):
        """Lowest eigenvalue of each Hamiltonian by Lanczos on the device (the energy of
        ``solve(ExactDiagonalizationProblem(H))``, ext/KrylovKitExt.jl:36-78), started from
        ``v0`` (default: the first Hamiltonian's start address).  Returns (energies, iterations)."""
        if v0 is None:
            v0 = self.unit_vector(self.hamiltonians[0].starting_address())
        v0 = self._vec(v0)
        e = np.zeros(self.R)
        it = ctypes.c_int64(0)
        _native.check(_native.lib().rimu_dense_ground_state(self._h, _p(v0), int(max_iters), float(tol),
                                                            _p(e), ctypes.byref(it)))
        return e, int(it.value)

    def unit_vector(self, address) -> np.ndarray:
        """The start vector e_address in every right-hand side."""
        b = self.basis()
        i = int(np.searchsorted(b, np.uint64(address.bits)))
        if i >= len(b) or b[i] != address.bits:
            raise ValueError("address not in the basis")
        v = np.zeros((self.dim, self.R))
        v[i, :] = 1.0
        return v
