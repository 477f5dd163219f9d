"""ctypes front end of the CPU oracle (``liboracle.so``) plus project-owned exact
diagonalisation.

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py``.  The product package never imports it.

Reference files restated here (paths relative to the Rimu.jl v0.14.0 checkout):

* ``HubbardMom1D`` momentum grid and dispersion: ``src/Hamiltonians/HubbardMom1D.jl:49-65``
  and ``:7`` (``hubbard_dispersion``).
* Basis by breadth-first search from a starting address:
  ``src/ExactDiagonalization/basis_breadth_first_search.jl:216,390``.
* Exact diagonalisation (``solve(ExactDiagonalizationProblem(H))``):
  ``src/ExactDiagonalization/solve.jl:72-112`` -- restated with scipy's dense ``eigh`` for
  small bases and ARPACK ``eigsh`` (Lanczos) for large ones.
"""
from __future__ import annotations

import ctypes
import math
import os
import pathlib
from fractions import Fraction

import numpy as np

_HERE = pathlib.Path(__file__).resolve().parent
_LIB = None

MODEL_REAL1D = 1
MODEL_MOM1D = 2
MODEL_REALSPACE = 3
STYLE_INTEGER = 1
STYLE_DETERMINISTIC = 2
STYLE_DYNAMIC = 3


class Model(ctypes.Structure):
    _fields_ = [("model", ctypes.c_int32), ("M", ctypes.c_int32), ("N", ctypes.c_int32),
                ("u", ctypes.c_double), ("t", ctypes.c_double), ("kes", ctypes.c_double * 64),
                ("ncomp", ctypes.c_int32), ("N2", ctypes.c_int32), ("fermi", ctypes.c_int32),
                ("ndir", ctypes.c_int32), ("has_u", ctypes.c_int32), ("has_pot", ctypes.c_int32),
                ("t2", ctypes.c_double), ("u12", ctypes.c_double), ("u22", ctypes.c_double),
                ("pot2", ctypes.c_double * 64), ("nbr", (ctypes.c_uint8 * 6) * 64)]


class Style(ctypes.Structure):
    _fields_ = [("style", ctypes.c_int32), ("compression", ctypes.c_int32),
                ("compression_threshold", ctypes.c_double), ("proj_threshold", ctypes.c_double),
                ("rel_threshold", ctypes.c_double), ("abs_threshold", ctypes.c_double)]


def lib():
    """Load ``liboracle.so`` (built by ``make -C oracle``)."""
    global _LIB
    if _LIB is None:
        # RIMU_ORACLE_LIB selects another build of the same source (the sanitizer build of
        # `make -C oracle asan`)
        path = pathlib.Path(os.environ["RIMU_ORACLE_LIB"]).resolve() if os.environ.get("RIMU_ORACLE_LIB") \
            else _HERE / "liboracle.so"
        if not path.exists():
            raise RuntimeError(f"oracle library missing: {path} (run `make -C oracle`)")
        L = ctypes.CDLL(str(path))
        P = ctypes.POINTER
        u64p, i32p, dp, vp = P(ctypes.c_uint64), P(ctypes.c_int32), P(ctypes.c_double), ctypes.c_void_p
        L.or_philox4x32_10.argtypes = [P(ctypes.c_uint32), P(ctypes.c_uint32), P(ctypes.c_uint32)]
        L.or_from_onr.argtypes = [i32p, ctypes.c_int32]
        L.or_from_onr.restype = ctypes.c_uint64
        L.or_to_onr.argtypes = [ctypes.c_uint64, ctypes.c_int32, i32p]
        L.or_occupied_modes.argtypes = [ctypes.c_uint64, i32p, i32p, i32p]
        L.or_occupied_modes.restype = ctypes.c_int32
        L.or_bose_hubbard_interaction.argtypes = [ctypes.c_uint64]
        L.or_bose_hubbard_interaction.restype = ctypes.c_int64
        L.or_num_offdiagonals.argtypes = [P(Model), ctypes.c_uint64]
        L.or_num_offdiagonals.restype = ctypes.c_int64
        L.or_diagonal_element.argtypes = [P(Model), ctypes.c_uint64]
        L.or_diagonal_element.restype = ctypes.c_double
        L.or_get_offdiagonal.argtypes = [P(Model), ctypes.c_uint64, ctypes.c_int64, dp]
        L.or_get_offdiagonal.restype = ctypes.c_uint64
        L.or_step.argtypes = [P(Model), P(Style), ctypes.c_uint64, ctypes.c_int64, ctypes.c_double,
                              ctypes.c_double, ctypes.c_double, ctypes.c_int32, ctypes.c_int32,
                              u64p, vp, ctypes.c_int64, u64p, vp, ctypes.c_int64, dp]
        L.or_step.restype = ctypes.c_int64
        L.or_step_initiator.argtypes = [P(Model), P(Style), ctypes.c_uint64, ctypes.c_int64,
                                        ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                        ctypes.c_double, u64p, vp, ctypes.c_int64, u64p, vp,
                                        ctypes.c_int64, dp]
        L.or_step_initiator.restype = ctypes.c_int64
        L.or_step_threaded.argtypes = [P(Model), P(Style), ctypes.c_uint64, ctypes.c_int64,
                                       ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_int32, ctypes.c_int32, u64p, vp, ctypes.c_int64,
                                       u64p, vp, ctypes.c_int64, dp]
        L.or_step_threaded.restype = ctypes.c_int64
        L.or_annihilate.argtypes = [ctypes.c_int32, u64p, vp, ctypes.c_int64, u64p, vp, ctypes.c_int64]
        L.or_annihilate.restype = ctypes.c_int64
        L.or_double_log_update.argtypes = [ctypes.c_double] * 7
        L.or_double_log_update.restype = ctypes.c_double
        L.or_build_basis.argtypes = [ctypes.c_int32, ctypes.c_int32, u64p, ctypes.c_int64]
        L.or_build_basis.restype = ctypes.c_int64
        L.or_apply_affine.argtypes = [P(Model), ctypes.c_int32, u64p, ctypes.c_int64, dp, dp, dp, dp]
        L.or_apply_affine.restype = ctypes.c_int32
        L.or_matrix_triplets.argtypes = [P(Model), u64p, ctypes.c_int64, P(ctypes.c_int64),
                                         P(ctypes.c_int64), dp, ctypes.c_int64]
        L.or_matrix_triplets.restype = ctypes.c_int64
        L.or_g2_real.argtypes = [ctypes.c_uint64, ctypes.c_int32, ctypes.c_int32]
        L.or_g2_real.restype = ctypes.c_double
        L.or_density_diagonal.argtypes = [ctypes.c_uint64, ctypes.c_int32, ctypes.c_int32]
        L.or_density_diagonal.restype = ctypes.c_double
        L.or_overlaps.argtypes = [u64p, vp, ctypes.c_int32, ctypes.c_int64, u64p, vp, ctypes.c_int32,
                                  ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                  i32p, i32p, P(Model), dp]
        u64 = ctypes.c_uint64
        L.or_from_onr2.argtypes = [i32p, ctypes.c_int32, u64p]
        L.or_to_onr2.argtypes = [u64, u64, ctypes.c_int32, i32p]
        L.or_num_offdiagonals2.argtypes = [P(Model), u64, u64]
        L.or_num_offdiagonals2.restype = ctypes.c_int64
        L.or_diagonal_element2.argtypes = [P(Model), u64, u64]
        L.or_diagonal_element2.restype = ctypes.c_double
        L.or_get_offdiagonal2.argtypes = [P(Model), u64, u64, ctypes.c_int64, dp, u64p]
        L.or_overlaps.restype = ctypes.c_int32
        L.or_rank.argtypes = [ctypes.c_uint64]
        L.or_rank.restype = ctypes.c_int64
        _LIB = L
    return _LIB


def _ptr(a, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


# ----------------------------------------------------------------------------------------
# Models
# ----------------------------------------------------------------------------------------
def mom1d_grid(M: int, t: float):
    """``ks`` and ``kes`` of ``HubbardMom1D`` (HubbardMom1D.jl:53-63).

    ``ks`` is Julia's ``range(start; step, length=M)`` on Float64 values: element j is the
    correctly rounded value of ``start + (j-1)*step``, computed here in exact rational
    arithmetic from the two Float64 inputs."""
    step = 2 * math.pi / M
    if M % 2 == 1:
        start = -math.pi * (1 + 1 / M) + step
    else:
        start = -math.pi + step
    ks = [float(Fraction(start) + j * Fraction(step)) for j in range(M)]
    kes = [(-2 * float(t)) * math.cos(k) for k in ks]
    return np.array(ks), np.array(kes)


def real1d(M: int, N: int, u: float = 1.0, t: float = 1.0) -> Model:
    m = Model()
    m.model, m.M, m.N, m.u, m.t = MODEL_REAL1D, M, N, float(u), float(t)
    return m


def mom1d(M: int, N: int, u: float = 1.0, t: float = 1.0) -> Model:
    m = Model()
    m.model, m.M, m.N, m.u, m.t = MODEL_MOM1D, M, N, float(u), float(t)
    _, kes = mom1d_grid(M, t)
    for j in range(M):
        m.kes[j] = kes[j]
    return m


def cubic_neighbours(dims, fold):
    """neighbor_site(geometry, s, d) - 1 for every site s and direction d, 255 where the move
    leaves a hard wall.  Sites are linear indices with the first dimension fastest
    (CartesianIndices / LinearIndices, geometry.jl:116-121); directions are +e_1..+e_D then
    -e_1..-e_D (Directions, geometry.jl:127-160); a periodic dimension folds with mod1
    (fold_vec, geometry.jl:100-112)."""
    D = len(dims)
    table = []
    for s in range(math.prod(dims)):
        x, r = [], s
        for n in dims:
            x.append(r % n + 1)
            r //= n
        row = []
        for d in range(2 * D):
            y = list(x)
            y[d % D] += 1 if d < D else -1
            ok = True
            for i in range(D):
                if fold[i]:
                    y[i] = (y[i] - 1) % dims[i] + 1
                elif not 1 <= y[i] <= dims[i]:
                    ok = False
            if not ok:
                row.append(255)
                continue
            lin, mul = 0, 1
            for i in range(D):
                lin += (y[i] - 1) * mul
                mul *= dims[i]
            row.append(lin)
        table.append(row)
    return table


def trap_potential(dims, v_c):
    """Potential per site of one component: sum_d v_c[d] * x_d^2 with x_d running over
    range(-fld(n_d, 2); length = n_d) (HubbardRealSpace.jl:216-223)."""
    out = []
    for s in range(math.prod(dims)):
        r, acc = s, None
        for d, n in enumerate(dims):
            x = (r % n) - n // 2
            r //= n
            term = v_c[d] * (x * x)
            acc = term if acc is None else acc + term
        out.append(acc)
    return out


def real_space(dims, comps, t=None, u=None, v=None, fold=None) -> Model:
    """HubbardRealSpace on a CubicGrid (PeriodicBoundaries by default).  comps: one or two
    ("bose" | "fermi", N) pairs; t: per component (default ones); u: C x C (default ones);
    v: C x D trap strengths (default zeros)."""
    C = len(comps)
    D = len(dims)
    fold = tuple(fold) if fold is not None else (True,) * D
    t = list(t) if t is not None else [1.0] * C
    u = np.array(u if u is not None else np.ones((C, C)), dtype=float).reshape(C, C)
    v = np.array(v if v is not None else np.zeros((C, D)), dtype=float).reshape(C, D)
    m = Model()
    m.model, m.M, m.N = MODEL_REALSPACE, math.prod(dims), comps[0][1]
    m.ncomp, m.N2 = C, comps[1][1] if C == 2 else 0
    m.fermi = sum(1 << c for c, (kind, _) in enumerate(comps) if kind == "fermi")
    m.ndir = 2 * D
    m.t, m.u = float(t[0]), float(u[0, 0])
    if C == 2:
        m.t2, m.u12, m.u22 = float(t[1]), float(u[0, 1]), float(u[1, 1])
    m.has_u = int(bool(np.any(u != 0)))
    m.has_pot = int(bool(np.any(v != 0)))
    if m.has_pot:
        for j, x in enumerate(trap_potential(dims, v[0])):
            m.kes[j] = x
        if C == 2:
            for j, x in enumerate(trap_potential(dims, v[1])):
                m.pot2[j] = x
    for s, row in enumerate(cubic_neighbours(dims, fold)):
        for d, x in enumerate(row):
            m.nbr[s][d] = x
    return m


def rs_key(model, onrs) -> int:
    """Key of a HubbardRealSpace address from the occupation numbers of each component."""
    parts = []
    for c, onr in enumerate(onrs):
        if (model.fermi >> c) & 1:
            parts.append((sum(1 << i for i, n in enumerate(onr) if n), model.M))
        else:
            parts.append((from_onr(onr) if sum(onr) else 0, sum(onr) + model.M - 1))
    key = 0
    for bits, w in parts:
        key = (key << w) | bits
    return key


def rs_onrs(model, key: int):
    """Occupation numbers of each component of a HubbardRealSpace key."""
    widths = [model.M if (model.fermi >> c) & 1 else n + model.M - 1
              for c, n in enumerate([model.N, model.N2][:model.ncomp])]
    out = []
    for c in reversed(range(model.ncomp)):
        w = widths[c]
        bits = key & ((1 << w) - 1)
        key >>= w
        if (model.fermi >> c) & 1:
            out.append(tuple((bits >> i) & 1 for i in range(model.M)))
        else:
            out.append(to_onr(bits, model.M))
    return tuple(reversed(out))


def style(kind: str, compression_threshold: float = 1.0, proj_threshold: float = 0.0,
          rel_threshold: float = 1.0, abs_threshold: float = math.inf, compression=None) -> Style:
    s = Style()
    s.style = {"integer": STYLE_INTEGER, "deterministic": STYLE_DETERMINISTIC,
               "dynamic": STYLE_DYNAMIC}[kind]
    if compression is None:
        compression = 1 if kind == "dynamic" else 0
    s.compression = compression
    s.compression_threshold = compression_threshold
    s.proj_threshold = proj_threshold
    s.rel_threshold = rel_threshold
    s.abs_threshold = abs_threshold
    return s


# ----------------------------------------------------------------------------------------
# Addresses and matrix elements
# ----------------------------------------------------------------------------------------
MASK64 = (1 << 64) - 1


def model_bits(model) -> int:
    """Bits of an address: N + M - 1 for the BoseFS models; for HubbardRealSpace the sum
    over components (M for a FermiFS, N_c + M - 1 for a BoseFS)."""
    if model.model != 3:
        return model.M + model.N - 1
    w = lambda c, n: model.M if (model.fermi >> c) & 1 else n + model.M - 1
    return w(0, model.N) + (w(1, model.N2) if model.ncomp == 2 else 0)


def key_words(model) -> int:
    """64-bit words per address: 1, or 2 when the address has more than 64 bits."""
    return 2 if model_bits(model) > 64 else 1


def pack_keys(ints, kw: int) -> np.ndarray:
    """Python-int addresses as uint64 words: shape (n,) for kw = 1, (n, 2) [lo, hi] for kw = 2."""
    ints = [int(x) for x in ints]
    if kw == 1:
        return np.array(ints, dtype=np.uint64)
    return np.array([[x & MASK64, x >> 64] for x in ints], dtype=np.uint64).reshape(len(ints), 2)


def unpack_keys(arr) -> list:
    arr = np.asarray(arr, dtype=np.uint64)
    if arr.ndim == 1:
        return [int(x) for x in arr.tolist()]
    return [int(lo) | (int(hi) << 64) for lo, hi in arr.tolist()]


def from_onr(onr) -> int:
    a = np.ascontiguousarray(onr, dtype=np.int32)
    if len(a) + int(a.sum()) - 1 > 64:
        out = np.zeros(2, dtype=np.uint64)
        lib().or_from_onr2(_ptr(a, ctypes.c_int32), len(a), _ptr(out, ctypes.c_uint64))
        return int(out[0]) | (int(out[1]) << 64)
    return int(lib().or_from_onr(_ptr(a, ctypes.c_int32), len(a)))


def to_onr(bits: int, M: int):
    out = np.zeros(M, dtype=np.int32)
    if bits >> 64:
        lib().or_to_onr2(bits & MASK64, bits >> 64, M, _ptr(out, ctypes.c_int32))
    else:
        lib().or_to_onr(ctypes.c_uint64(bits), M, _ptr(out, ctypes.c_int32))
    return tuple(int(x) for x in out)


def occupied_modes(bits: int):
    o, m, f = (np.zeros(64, dtype=np.int32) for _ in range(3))
    n = lib().or_occupied_modes(ctypes.c_uint64(bits), _ptr(o, ctypes.c_int32),
                                _ptr(m, ctypes.c_int32), _ptr(f, ctypes.c_int32))
    return [(int(o[i]), int(m[i]), int(f[i])) for i in range(n)]


def bose_hubbard_interaction(bits: int) -> int:
    return int(lib().or_bose_hubbard_interaction(ctypes.c_uint64(bits)))


def num_offdiagonals(model: Model, bits: int) -> int:
    if key_words(model) == 2:
        return int(lib().or_num_offdiagonals2(ctypes.byref(model), bits & MASK64, bits >> 64))
    return int(lib().or_num_offdiagonals(ctypes.byref(model), ctypes.c_uint64(bits)))


def diagonal_element(model: Model, bits: int) -> float:
    if key_words(model) == 2:
        return float(lib().or_diagonal_element2(ctypes.byref(model), bits & MASK64, bits >> 64))
    return float(lib().or_diagonal_element(ctypes.byref(model), ctypes.c_uint64(bits)))


def get_offdiagonal(model: Model, bits: int, chosen: int):
    v = ctypes.c_double()
    if key_words(model) == 2:
        out = np.zeros(2, dtype=np.uint64)
        lib().or_get_offdiagonal2(ctypes.byref(model), bits & MASK64, bits >> 64, chosen,
                                  ctypes.byref(v), _ptr(out, ctypes.c_uint64))
        return int(out[0]) | (int(out[1]) << 64), float(v.value)
    nb = lib().or_get_offdiagonal(ctypes.byref(model), ctypes.c_uint64(bits), chosen, ctypes.byref(v))
    return int(nb), float(v.value)


def offdiagonals(model: Model, bits: int):
    return [get_offdiagonal(model, bits, c) for c in range(1, num_offdiagonals(model, bits) + 1)]


def philox4x32_10(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib().or_philox4x32_10(c, k, o)
    return tuple(int(x) for x in o)


# ----------------------------------------------------------------------------------------
# Steps and annihilation
# ----------------------------------------------------------------------------------------
STAT_NAMES = ("attempts", "exact_steps", "inexact_steps", "spawns", "deaths", "clones",
              "zombies", "len_before", "len", "norm")


def step(model: Model, sty: Style, keys, vals, *, seed: int, step: int, shift: float,
         dtau: float, boost: float = 1.0, order: int = 0, threads: int = 0):
    """One ``apply_operator!`` + ``compress!`` of (keys, vals).  ``threads > 0`` uses the
    threaded PDVec-style restatement (output unsorted, one-chunk addresses only).  Keys are
    (n,) uint64, or (n, 2) [lo, hi] words for two-chunk addresses."""
    kw = key_words(model)
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    if kw == 2:
        keys = keys.reshape(-1, 2)
    vtype = 0 if sty.style == STYLE_INTEGER else 1
    vals = np.ascontiguousarray(vals, dtype=np.int64 if vtype == 0 else np.float64)
    cap = max(64, 4 * len(keys) + 1024)
    while True:
        ok = np.zeros((cap, 2) if kw == 2 else cap, dtype=np.uint64)
        ov = np.zeros(cap, dtype=vals.dtype)
        st = np.zeros(16, dtype=np.float64)
        if threads:
            n = lib().or_step_threaded(ctypes.byref(model), ctypes.byref(sty), seed, step, shift,
                                       dtau, boost, vtype, threads, _ptr(keys, ctypes.c_uint64),
                                       vals.ctypes.data, len(keys), _ptr(ok, ctypes.c_uint64),
                                       ov.ctypes.data, cap, _ptr(st, ctypes.c_double))
        else:
            n = lib().or_step(ctypes.byref(model), ctypes.byref(sty), seed, step, shift, dtau,
                              boost, order, vtype, _ptr(keys, ctypes.c_uint64), vals.ctypes.data,
                              len(keys), _ptr(ok, ctypes.c_uint64), ov.ctypes.data, cap,
                              _ptr(st, ctypes.c_double))
        if n >= 0:
            return ok[:n], ov[:n], dict(zip(STAT_NAMES, st[:10].tolist()))
        cap = -n + 16


INITIATOR_RULES = {"none": 0, "initiator": 1, "simple": 2, "coherent": 3}


def step_initiator(model: Model, sty: Style, keys, vals, *, rule: str, threshold: float,
                   seed: int, step: int, shift: float, dtau: float, boost: float = 1.0,
                   order: int = 0):
    """One step with an initiator rule (``rule`` in INITIATOR_RULES).  With NoCompression
    the result may hold zero values (addresses that received only discarded deposits)."""
    kw = key_words(model)
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    if kw == 2:
        keys = keys.reshape(-1, 2)
    vtype = 0 if sty.style == STYLE_INTEGER else 1
    vals = np.ascontiguousarray(vals, dtype=np.int64 if vtype == 0 else np.float64)
    cap = max(64, 4 * len(keys) + 1024)
    while True:
        ok = np.zeros((cap, 2) if kw == 2 else cap, dtype=np.uint64)
        ov = np.zeros(cap, dtype=vals.dtype)
        st = np.zeros(16, dtype=np.float64)
        n = lib().or_step_initiator(ctypes.byref(model), ctypes.byref(sty), seed, step, shift,
                                    dtau, boost, order, vtype, INITIATOR_RULES[rule],
                                    float(threshold), _ptr(keys, ctypes.c_uint64),
                                    vals.ctypes.data, len(keys), _ptr(ok, ctypes.c_uint64),
                                    ov.ctypes.data, cap, _ptr(st, ctypes.c_double))
        if n >= 0:
            return ok[:n], ov[:n], dict(zip(STAT_NAMES, st[:10].tolist()))
        cap = -n + 16


def annihilate(keys, vals):
    """Left fold of ``deposit!`` over a spawn list; output sorted by key."""
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    vals = np.ascontiguousarray(vals)
    vtype = 0 if vals.dtype == np.int64 else 1
    if vtype == 1:
        vals = vals.astype(np.float64, copy=False)
    cap = len(keys) + 1
    ok = np.zeros(cap, dtype=np.uint64)
    ov = np.zeros(cap, dtype=vals.dtype)
    n = lib().or_annihilate(vtype, _ptr(keys, ctypes.c_uint64), vals.ctypes.data, len(keys),
                            _ptr(ok, ctypes.c_uint64), ov.ctypes.data, cap)
    return ok[:n], ov[:n]


def g2_real(bits: int, M: int, d: int) -> float:
    return float(lib().or_g2_real(ctypes.c_uint64(bits), M, d))


def density_diagonal(bits: int, M: int, mode: int) -> float:
    return float(lib().or_density_diagonal(ctypes.c_uint64(bits), M, mode))


def overlaps(tkeys, tvals, ukeys, uvals, M, ops=()):
    """[dot(t, u), dot(t, op, u) for op in ops]; an op is ("ham", Model), ("g2", d) or
    ("density", mode)."""
    tk = np.ascontiguousarray(tkeys, dtype=np.uint64)
    uk = np.ascontiguousarray(ukeys, dtype=np.uint64)
    kw = 2 if tk.ndim == 2 or uk.ndim == 2 else 1
    tv = np.ascontiguousarray(tvals)
    uv = np.ascontiguousarray(uvals)
    tvt = 0 if tv.dtype == np.int64 else 1
    uvt = 0 if uv.dtype == np.int64 else 1
    tv = tv if tvt == 0 else tv.astype(np.float64)
    uv = uv if uvt == 0 else uv.astype(np.float64)
    n = len(ops)
    kinds = np.zeros(max(n, 1), dtype=np.int32)
    params = np.zeros(max(n, 1), dtype=np.int32)
    models = (Model * max(n, 1))()
    for i, (kind, arg) in enumerate(ops):
        if kind == "ham":
            kinds[i] = 1
            models[i] = arg
        elif kind == "g2":
            kinds[i], params[i] = 2, arg
        else:
            kinds[i], params[i] = 3, arg
    out = np.zeros(n + 1)
    lib().or_overlaps(_ptr(tk, ctypes.c_uint64), tv.ctypes.data, tvt, len(tk),
                      _ptr(uk, ctypes.c_uint64), uv.ctypes.data, uvt, len(uk), M, kw, n,
                      _ptr(kinds, ctypes.c_int32), _ptr(params, ctypes.c_int32), models,
                      _ptr(out, ctypes.c_double))
    return out


def double_log_update(shift, tnorm, pnorm, target, zeta=0.08, xi=None, dtau=0.01):
    xi = zeta ** 2 / 4 if xi is None else xi
    return float(lib().or_double_log_update(shift, tnorm, pnorm, target, zeta, xi, dtau))


# ----------------------------------------------------------------------------------------
# Bases, deterministic products and exact diagonalisation
# ----------------------------------------------------------------------------------------
def build_basis(M: int, N: int) -> np.ndarray:
    """All N-boson M-mode addresses, ascending."""
    cap = math.comb(N + M - 1, N)
    out = np.zeros(cap, dtype=np.uint64)
    n = lib().or_build_basis(M, N, _ptr(out, ctypes.c_uint64), cap)
    assert n == cap
    return out


def rank(bits: int) -> int:
    return int(lib().or_rank(ctypes.c_uint64(bits)))


def apply_affine(models, basis: np.ndarray, a, b, v: np.ndarray) -> np.ndarray:
    """w_r = (a_r + b_r H_r) v_r for nrhs models; v has shape (D, nrhs)."""
    models = list(models)
    nrhs = len(models)
    arr = (Model * nrhs)(*models)
    basis = np.ascontiguousarray(basis, dtype=np.uint64)
    v = np.ascontiguousarray(v, dtype=np.float64).reshape(len(basis), nrhs)
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    w = np.zeros_like(v)
    rc = lib().or_apply_affine(arr, nrhs, _ptr(basis, ctypes.c_uint64), len(basis),
                               _ptr(a, ctypes.c_double), _ptr(b, ctypes.c_double),
                               _ptr(v, ctypes.c_double), _ptr(w, ctypes.c_double))
    if rc != 0:
        raise RuntimeError("neighbour outside basis")
    return w


def bfs_basis(model: Model, start: int, limit: int = 5_000_000) -> np.ndarray:
    """Addresses connected to ``start`` by nonzero off-diagonal elements, ascending."""
    seen = {start}
    frontier = [start]
    while frontier:
        nxt = []
        for a in frontier:
            for nb, val in offdiagonals(model, a):
                if val != 0 and nb not in seen:
                    seen.add(nb)
                    nxt.append(nb)
        if len(seen) > limit:
            raise RuntimeError("basis too large")
        frontier = nxt
    return np.array(sorted(seen), dtype=np.uint64)


def sparse_hamiltonian(model: Model, basis: np.ndarray):
    """H on ``basis`` as a scipy CSR matrix (duplicate elements summed)."""
    import scipy.sparse as sp
    basis = np.ascontiguousarray(basis, dtype=np.uint64)
    D = len(basis)
    cap = D * 64
    while True:
        r = np.zeros(cap, dtype=np.int64)
        c = np.zeros(cap, dtype=np.int64)
        v = np.zeros(cap)
        n = lib().or_matrix_triplets(ctypes.byref(model), _ptr(basis, ctypes.c_uint64), D,
                                     _ptr(r, ctypes.c_int64), _ptr(c, ctypes.c_int64),
                                     _ptr(v, ctypes.c_double), cap)
        if n == -1 - cap:
            raise RuntimeError("neighbour outside basis")
        if n >= 0:
            return sp.csr_matrix((v[:n], (r[:n], c[:n])), shape=(D, D))
        cap = -n


def hamiltonian_operator(model: Model, basis: np.ndarray):
    return sparse_hamiltonian(model, basis)


def ed_energies(model: Model, basis: np.ndarray, k: int = 1) -> np.ndarray:
    """Lowest ``k`` eigenvalues of H on ``basis`` (dense ``eigh`` when small, else ARPACK)."""
    D = len(basis)
    if D <= 3000:
        H = sparse_hamiltonian(model, basis).toarray()
        return np.linalg.eigvalsh(0.5 * (H + H.T))[:k]
    import scipy.sparse.linalg as sla
    vals = sla.eigsh(sparse_hamiltonian(model, basis), k=k, which="SA", tol=1e-12,
                     return_eigenvectors=False)
    return np.sort(vals)
