"""ctypes binding of ``librimu_hip.so`` (declarations in ``include/rimu_hip.h``).

The library is the only compute path: nothing here falls back to the CPU.  Loading fails
loudly when the library has not been built, and every call that needs a device raises
``RimuError`` when there is none.
"""
from __future__ import annotations

import ctypes
import pathlib

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "librimu_hip.so"

RIMU_OK, RIMU_ERR_ARG, RIMU_ERR_HIP, RIMU_ERR_CAPACITY, RIMU_ERR_NODEVICE, RIMU_ERR_INTERNAL = range(6)
MODEL_REAL1D, MODEL_MOM1D, MODEL_REALSPACE = 1, 2, 3
STYLE_INTEGER, STYLE_DETERMINISTIC, STYLE_DYNAMIC = 1, 2, 3
INITIATOR_NONE, INITIATOR_DEFAULT, INITIATOR_SIMPLE, INITIATOR_COHERENT = 0, 1, 2, 3
OP_HAMILTONIAN, OP_G2_REAL, OP_DENSITY_DIAGONAL = 1, 2, 3
SHIFT_HOST, SHIFT_DOUBLE_LOG = 0, 1
MAX_OPERATORS = 8

# Every symbol declared in include/rimu_hip.h (checked by the CPU test-suite).
EXPORTS = (
    "rimu_last_error", "rimu_device_count", "rimu_engine_create", "rimu_engine_destroy",
    "rimu_engine_reserve", "rimu_vector_upload", "rimu_vector_length", "rimu_vector_download",
    "rimu_set_projector", "rimu_step", "rimu_engine_set_shift_update", "rimu_engine_get_shift",
    "rimu_engine_set_initiator", "rimu_overlaps",
    "rimu_engine_set_sort_bits", "rimu_engine_set_radix_mode", "rimu_engine_set_lookback",
    "rimu_engine_set_timing",
    "rimu_engine_kernel_times",
    "rimu_engine_last_traffic", "rimu_annihilate", "rimu_annihilate_noalloc", "rimu_hamiltonian_eval",
    "rimu_basis_size", "rimu_dense_create", "rimu_dense_destroy", "rimu_dense_dim",
    "rimu_dense_basis", "rimu_dense_apply", "rimu_dense_pullback", "rimu_dense_projector",
    "rimu_dense_ground_state",
    "rimu_batch_create", "rimu_batch_destroy", "rimu_batch_set_config", "rimu_batch_set_persistent",
    "rimu_batch_set_wait", "rimu_batch_step",
    "rimu_batch_history", "rimu_batch_status", "rimu_batch_autotune", "rimu_batch_vector",
)


class RimuError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[rimu code {code}] {msg}")
        self.code = code


class CapacityError(RimuError):
    pass


class Model(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("M", ctypes.c_int32), ("N", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("u", ctypes.c_double), ("t", ctypes.c_double),
                ("kes", ctypes.c_double * 64),
                ("ncomp", ctypes.c_int32), ("N2", ctypes.c_int32), ("fermi", ctypes.c_int32),
                ("ndir", ctypes.c_int32), ("has_u", ctypes.c_int32), ("has_pot", ctypes.c_int32),
                ("t2", ctypes.c_double), ("u12", ctypes.c_double), ("u22", ctypes.c_double),
                ("pot2", ctypes.c_double * 64), ("nbr", (ctypes.c_uint8 * 6) * 64)]


class Style(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("compression", ctypes.c_int32),
                ("compression_threshold", ctypes.c_double), ("proj_threshold", ctypes.c_double),
                ("rel_threshold", ctypes.c_double), ("abs_threshold", ctypes.c_double)]


class Operator(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("param", ctypes.c_int32), ("model", Model)]


class BatchRun(ctypes.Structure):
    _fields_ = [("M", ctypes.c_int32), ("N", ctypes.c_int32), ("u", ctypes.c_double),
                ("t", ctypes.c_double), ("seed", ctypes.c_uint64), ("start_address", ctypes.c_uint64),
                ("start_population", ctypes.c_int64), ("shift", ctypes.c_double),
                ("target_walkers", ctypes.c_double), ("zeta", ctypes.c_double),
                ("xi", ctypes.c_double), ("dtau", ctypes.c_double), ("max_length", ctypes.c_int64)]


class StepStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "spawn_attempts", "exact_steps", "inexact_steps", "spawns", "deaths", "clones",
        "zombies", "len_before", "len", "norm", "vproj", "hproj")] + [
        ("overflow", ctypes.c_int32), ("reserved", ctypes.c_int32)]


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise RimuError(RIMU_ERR_INTERNAL,
                            f"{LIB_PATH} is not built; run `python rimu.jl_amd/build.py`")
        L = ctypes.CDLL(str(LIB_PATH))
        P = ctypes.POINTER
        vp, i64, u64, i32, dbl = ctypes.c_void_p, ctypes.c_int64, ctypes.c_uint64, ctypes.c_int32, ctypes.c_double
        L.rimu_last_error.restype = ctypes.c_char_p
        L.rimu_device_count.argtypes = [P(i32)]
        L.rimu_engine_create.argtypes = [P(Model), P(Style), u64, i64, i64, P(vp)]
        L.rimu_engine_destroy.argtypes = [vp]
        L.rimu_engine_reserve.argtypes = [vp, i64, i64]
        L.rimu_vector_upload.argtypes = [vp, vp, vp, i64]
        L.rimu_vector_length.argtypes = [vp, P(i64)]
        L.rimu_vector_download.argtypes = [vp, vp, vp, i64, P(i64)]
        L.rimu_set_projector.argtypes = [vp, vp, vp, i64, vp, vp, i64]
        L.rimu_step.argtypes = [vp, i64, dbl, dbl, dbl, P(StepStats)]
        L.rimu_engine_set_shift_update.argtypes = [vp, i32, dbl, dbl, dbl, dbl, dbl]
        L.rimu_engine_get_shift.argtypes = [vp, P(dbl), P(dbl)]
        L.rimu_engine_set_initiator.argtypes = [vp, i32, dbl]
        L.rimu_engine_set_sort_bits.argtypes = [vp, i32]
        L.rimu_engine_set_radix_mode.argtypes = [vp, i32]
        L.rimu_engine_set_lookback.argtypes = [vp, ctypes.c_uint32, i32]
        L.rimu_overlaps.argtypes = [vp, vp, P(Operator), i32, P(dbl)]
        L.rimu_engine_set_timing.argtypes = [vp, i32]
        L.rimu_engine_kernel_times.argtypes = [vp, ctypes.c_char_p, i32, P(dbl), P(i64), i32, P(i32)]
        L.rimu_engine_last_traffic.argtypes = [vp, P(dbl), P(dbl)]
        L.rimu_annihilate.argtypes = [i32, vp, vp, i64, vp, vp, i64, P(i64)]
        L.rimu_annihilate_noalloc.argtypes = [i32, vp, vp, i64, vp, vp, i64, P(i64)]
        L.rimu_hamiltonian_eval.argtypes = [P(Model), vp, vp, i64, vp, vp, vp, vp]
        L.rimu_basis_size.argtypes = [i32, i32, P(i64)]
        L.rimu_dense_create.argtypes = [P(Model), i32, P(vp)]
        L.rimu_dense_destroy.argtypes = [vp]
        L.rimu_dense_dim.argtypes = [vp, P(i64)]
        L.rimu_dense_basis.argtypes = [vp, vp]
        L.rimu_dense_apply.argtypes = [vp, vp, vp, vp, vp]
        L.rimu_dense_pullback.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp]
        L.rimu_dense_projector.argtypes = [vp, i64, dbl, vp, i64, vp, vp, vp, vp, vp, vp, vp]
        L.rimu_dense_ground_state.argtypes = [vp, vp, i64, dbl, vp, P(i64)]
        L.rimu_batch_create.argtypes = [P(BatchRun), i32, i64, P(vp)]
        L.rimu_batch_destroy.argtypes = [vp]
        L.rimu_batch_set_config.argtypes = [vp, i32, i32]
        L.rimu_batch_set_persistent.argtypes = [vp, i32]
        L.rimu_batch_set_wait.argtypes = [vp, ctypes.c_uint32, i32]
        L.rimu_batch_step.argtypes = [vp, i64, i32]
        L.rimu_batch_history.argtypes = [vp, P(i64), vp, vp, vp, vp, vp, vp]
        L.rimu_batch_status.argtypes = [vp, vp, vp]
        L.rimu_batch_autotune.argtypes = [vp, i64, P(i32), P(i32), P(i32), vp]
        L.rimu_batch_vector.argtypes = [vp, i32, vp, vp, i64, P(i64)]
        for name in EXPORTS:
            if name != "rimu_last_error":
                getattr(L, name).restype = ctypes.c_int
        _LIB = L
    return _LIB


def check(rc: int):
    if rc != RIMU_OK:
        msg = lib().rimu_last_error().decode(errors="replace")
        if rc == RIMU_ERR_CAPACITY:
            raise CapacityError(rc, msg)
        raise RimuError(rc, msg)


def device_count() -> int:
    n = ctypes.c_int32(0)
    check(lib().rimu_device_count(ctypes.byref(n)))
    return int(n.value)


def _p(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data)


def annihilate(keys, vals, noalloc: bool = False):
    """Device annihilation of a spawn list (``rimu_annihilate``; with ``noalloc``,
    ``rimu_annihilate_noalloc``, which takes at most 2^20 records and allocates no device
    memory)."""
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    vals = np.ascontiguousarray(vals)
    vtype = 0 if vals.dtype.kind in "iu" else 1
    vals = vals.astype(np.int64 if vtype == 0 else np.float64, copy=False)
    n = len(keys)
    ok = np.zeros(max(n, 1), dtype=np.uint64)
    ov = np.zeros(max(n, 1), dtype=vals.dtype)
    nout = ctypes.c_int64(0)
    fn = lib().rimu_annihilate_noalloc if noalloc else lib().rimu_annihilate
    check(fn(vtype, _p(keys), _p(vals), n, _p(ok), _p(ov), max(n, 1), ctypes.byref(nout)))
    m = nout.value
    return ok[:m].copy(), ov[:m].copy()


def overlaps(left: "Engine", right: "Engine", ops=()) -> np.ndarray:
    """[dot(left, right), dot(left, op, right) for op in ops] (``rimu_overlaps``); each op is
    an ``Operator`` structure."""
    n = len(ops)
    if n > MAX_OPERATORS:
        raise ValueError(f"at most {MAX_OPERATORS} operators per call")
    arr = (Operator * max(n, 1))(*ops)
    out = (ctypes.c_double * (n + 1))()
    check(lib().rimu_overlaps(left._h, right._h, arr, n, out))
    return np.array(out[:], dtype=np.float64)


def model_key_bits(model: Model) -> int:
    """Bits of an address of the model (N + M - 1 for BoseFS; for HubbardRealSpace the sum
    over components of M (FermiFS) or N_c + M - 1 (BoseFS))."""
    if model.kind != MODEL_REALSPACE:
        return model.M + model.N - 1
    width = lambda c, n: model.M if (model.fermi >> c) & 1 else n + model.M - 1
    return width(0, model.N) + (width(1, model.N2) if model.ncomp == 2 else 0)


def model_key_words(model: Model) -> int:
    return 2 if model_key_bits(model) > 64 else 1


def hamiltonian_eval(model: Model, addrs, chosen):
    """Diagonal, number of off-diagonals and one chosen off-diagonal per address; addresses
    are word arrays ((n,) or (n, 2))."""
    kw = model_key_words(model)
    addrs = np.ascontiguousarray(addrs, dtype=np.uint64)
    if kw == 2:
        addrs = addrs.reshape(-1, 2)
    chosen = np.ascontiguousarray(chosen, dtype=np.int64)
    n = len(addrs)
    diag = np.zeros(max(n, 1))
    nod = np.zeros(max(n, 1), dtype=np.int64)
    na = np.zeros((max(n, 1), 2) if kw == 2 else max(n, 1), dtype=np.uint64)
    ov = np.zeros(max(n, 1))
    check(lib().rimu_hamiltonian_eval(ctypes.byref(model), _p(addrs), _p(chosen), n, _p(diag),
                                      _p(nod), _p(na), _p(ov)))
    return diag[:n], nod[:n], na[:n], ov[:n]


class Engine:
    """Owner of one ``rimu_engine`` handle: a device-resident vector plus step memory."""

    def __init__(self, model: Model, style: Style, seed: int, max_entries: int, max_spawns: int):
        self._h = ctypes.c_void_p()
        self.model = model
        self.style = style
        self.int_vals = style.kind == STYLE_INTEGER
        self.kw = model_key_words(model)
        self.max_entries = int(max_entries)
        self.max_spawns = int(max_spawns)
        check(lib().rimu_engine_create(ctypes.byref(model), ctypes.byref(style),
                                       ctypes.c_uint64(seed & (2 ** 64 - 1)), self.max_entries,
                                       self.max_spawns, ctypes.byref(self._h)))

    def close(self):
        if self._h:
            lib().rimu_engine_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def vdtype(self):
        return np.int64 if self.int_vals else np.float64

    def reserve(self, max_entries: int, max_spawns: int):
        check(lib().rimu_engine_reserve(self._h, int(max_entries), int(max_spawns)))
        self.max_entries = max(self.max_entries, int(max_entries))
        self.max_spawns = max(self.max_spawns, int(max_spawns))

    def upload(self, keys, vals):
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        if self.kw == 2:
            keys = keys.reshape(-1, 2)
            ascending = len(keys) < 2 or bool(np.all(
                (keys[1:, 1] > keys[:-1, 1]) | ((keys[1:, 1] == keys[:-1, 1]) & (keys[1:, 0] > keys[:-1, 0]))))
        else:
            ascending = len(keys) < 2 or bool(np.all(keys[1:] > keys[:-1]))
        vals = np.ascontiguousarray(vals, dtype=self.vdtype)
        if not ascending:
            order = np.argsort(keys, kind="stable") if self.kw == 1 else np.lexsort((keys[:, 0], keys[:, 1]))
            keys, vals = np.ascontiguousarray(keys[order]), vals[order]
            same = (keys[1:] == keys[:-1]) if self.kw == 1 else np.all(keys[1:] == keys[:-1], axis=1)
            if np.any(same):
                raise ValueError("duplicate keys in vector")
        check(lib().rimu_vector_upload(self._h, _p(keys), _p(vals), len(keys)))

    def __len__(self):
        n = ctypes.c_int64(0)
        check(lib().rimu_vector_length(self._h, ctypes.byref(n)))
        return int(n.value)

    def download(self):
        n = len(self)
        keys = np.zeros((max(n, 1), 2) if self.kw == 2 else max(n, 1), dtype=np.uint64)
        vals = np.zeros(max(n, 1), dtype=self.vdtype)
        out = ctypes.c_int64(0)
        check(lib().rimu_vector_download(self._h, _p(keys), _p(vals), max(n, 1), ctypes.byref(out)))
        return keys[:out.value], vals[:out.value]

    def set_projector(self, pk, pc, hk, hc):
        pk = np.ascontiguousarray(pk, dtype=np.uint64)
        pc = np.ascontiguousarray(pc, dtype=np.float64)
        hk = np.ascontiguousarray(hk, dtype=np.uint64)
        hc = np.ascontiguousarray(hc, dtype=np.float64)
        check(lib().rimu_set_projector(self._h, _p(pk), _p(pc), len(pk), _p(hk), _p(hc), len(hk)))

    def step(self, step: int, shift: float, dtau: float, boost: float = 1.0) -> StepStats:
        """One step; grows the spawn buffer and repeats the step when it overflows."""
        st = StepStats()
        while True:
            rc = lib().rimu_step(self._h, int(step), float(shift), float(dtau), float(boost),
                                 ctypes.byref(st))
            if rc == RIMU_ERR_CAPACITY and "spawn" in lib().rimu_last_error().decode():
                self.reserve(self.max_entries, min(2 * self.max_spawns + 1024, 2 ** 32 - 2))
                continue
            check(rc)
            return st

    def set_shift_update(self, mode: int, shift: float = 0.0, pnorm: float = 1.0,
                         target_walkers: float = 1000.0, zeta: float = 0.08, xi: float | None = None):
        """Keep the shift on the device and update it there after every step
        (``SHIFT_DOUBLE_LOG``: DoubleLogUpdate from the initial shift and pnorm), or return it to
        the caller (``SHIFT_HOST``).  ``step`` then ignores its shift argument."""
        xi = zeta ** 2 / 4 if xi is None else xi
        check(lib().rimu_engine_set_shift_update(self._h, int(mode), float(shift), float(pnorm),
                                                 float(target_walkers), float(zeta), float(xi)))

    def shift_parameters(self):
        """(shift, pnorm) of the device shift update after the last step."""
        s, n = ctypes.c_double(0), ctypes.c_double(0)
        check(lib().rimu_engine_get_shift(self._h, ctypes.byref(s), ctypes.byref(n)))
        return float(s.value), float(n.value)

    def set_sort_bits(self, bits: int):
        """Top bits of mixed keys that the radix passes cover (tests and tuning)."""
        check(lib().rimu_engine_set_sort_bits(self._h, int(bits)))

    def set_radix_mode(self, mode: int):
        """Sort and merge of the spawn records (``rimu_engine_set_radix_mode``): 2 top-bit sort
        and bucket merge where eligible (default), 1 look-back scatters over every bit and the
        merge path, 0 per-pass counts and the merge path."""
        check(lib().rimu_engine_set_radix_mode(self._h, int(mode)))

    def set_lookback(self, spin_limit: int, withhold_first: bool = False):
        """Bound of the look-back waits and the host-forced give-up test
        (``rimu_engine_set_lookback``)."""
        check(lib().rimu_engine_set_lookback(self._h, int(spin_limit), int(bool(withhold_first))))

    def set_initiator(self, rule: int, threshold: float = 1.0):
        """Initiator rule of the vector (INITIATOR_* codes) and its threshold."""
        check(lib().rimu_engine_set_initiator(self._h, int(rule), float(threshold)))

    def set_timing(self, enabled: bool):
        check(lib().rimu_engine_set_timing(self._h, 1 if enabled else 0))

    def kernel_times(self):
        cap, nl = 64, 64
        names = ctypes.create_string_buffer(cap * nl)
        ms = (ctypes.c_double * cap)()
        launches = (ctypes.c_int64 * cap)()
        n = ctypes.c_int32(0)
        check(lib().rimu_engine_kernel_times(self._h, names, nl, ms, launches, cap, ctypes.byref(n)))
        out = {}
        for i in range(n.value):
            nm = names.raw[i * nl:(i + 1) * nl].split(b"\0", 1)[0].decode()
            out[nm] = (float(ms[i]), int(launches[i]))
        return out

    def last_traffic(self):
        b, r = ctypes.c_double(0), ctypes.c_double(0)
        check(lib().rimu_engine_last_traffic(self._h, ctypes.byref(b), ctypes.byref(r)))
        return float(b.value), float(r.value)
