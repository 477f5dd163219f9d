"""``save_state`` / ``load_state``: Rimu's Arrow state files (``src/RimuIO/RimuIO.jl:92-180``)
and ``save_df`` / ``load_df`` for reports (``:34-79``).

Layout of a state file, as ``Arrow.write(filename, Tables.table(vector); compress=:zstd,
metadata)`` produces it (``RimuIO.jl:113-118``, the table of ``tables.jl``):

* Arrow IPC file format, zstd-compressed record batches;
* two non-nullable columns ``key`` and ``value`` (``load_state`` refuses any other schema,
  ``RimuIO.jl:149-152``);
* ``key``: the Arrow extension type of BoseFS on a BitString (``arrowtypes.jl:58-83``):
  extension name ``Rimu.BoseFS.BitString``, extension metadata ``"N.M.B"`` (particles,
  modes, bits), storage ``NTuple{K,T}`` = FixedSizeList of K non-nullable chunks, where the
  chunk type T is UInt8/16/32/64 for B <= 8/16/32/64 and K = 2 UInt64 chunks for
  64 < B <= 128 (``num_chunks``, ``bitstring.jl:6-18``); chunk 1 holds the most significant
  bits (``bitstring.jl:64-75``, mode 1 in the lowest bits of the last chunk);
* ``value``: Int64 (integer walkers) or Float64;
* schema metadata: ``RIMU_PACKAGE_VERSION`` followed by the keyword arguments of
  ``save_state``, each as a string; ``load_state`` parses each back as Int, Float64,
  ComplexF64, Bool or leaves it a string (``RimuIO.jl:158-173``).

The Arrow conventions follow Arrow.jl (Rimu's dependency, compat "1.5, 2"): an extension
type is the field metadata pair ``ARROW:extension:name`` / ``ARROW:extension:metadata``,
an ``NTuple{N,T}`` is a FixedSizeList<T>[N], and a field is nullable only for a
``Union{Missing,...}`` element type.
"""
from __future__ import annotations

import re

import numpy as np

from .addresses import BoseFS, key_words, pack_keys, unpack_keys
from .dictvectors import DVec
from .styles import IsDynamicSemistochastic, IsStochasticInteger

RIMU_PACKAGE_VERSION = "0.14.0"   # the Rimu.jl release whose format this mirrors
BOSEFS_BS = "Rimu.BoseFS.BitString"

_CHUNK_TYPES = ((8, np.uint8), (16, np.uint16), (32, np.uint32), (64, np.uint64))


def _pa():
    import pyarrow as pa
    import pyarrow.ipc as ipc
    return pa, ipc


def chunk_dtype(B: int):
    """Chunk type of a BitString of B bits (num_chunks, bitstring.jl:6-18)."""
    for bits, dt in _CHUNK_TYPES:
        if B <= bits:
            return dt
    if B <= 128:
        return np.uint64
    raise ValueError("addresses above 128 bits are not supported")


def _julia_string(v) -> str:
    """``string(v)`` of the Julia value a Python keyword stands for."""
    if isinstance(v, bool):
        return "true" if v else "false"
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    if isinstance(v, (float, np.floating)):
        return repr(float(v))
    if isinstance(v, complex):
        return f"{repr(v.real)} + {repr(v.imag)}im" if v.imag >= 0 or np.isnan(v.imag) \
            else f"{repr(v.real)} - {repr(-v.imag)}im"
    return str(v)


_COMPLEX_RE = re.compile(r"^\s*([^\s]+)\s*([+-])\s*([^\s]+)im\s*$")


def _parse_meta(k: str, v: str):
    """``load_state``'s metadata parse (RimuIO.jl:158-173)."""
    if k == "RIMU_PACKAGE_VERSION":
        return v
    try:
        if re.fullmatch(r"[+-]?\d+", v.strip()):
            return int(v)
    except ValueError:
        pass
    try:
        return float(v)
    except ValueError:
        pass
    m = _COMPLEX_RE.match(v)
    if m:
        try:
            re_, im_ = float(m.group(1)), float(m.group(3))
            return complex(re_, im_ if m.group(2) == "+" else -im_)
        except ValueError:
            pass
    if v in ("true", "false"):
        return v == "true"
    return v


def _vector_arrays(vector):
    """(M, N, keys u64, values) of a DVec, a DeviceDVec or a simulation's state vector."""
    if hasattr(vector, "state_vector"):
        vector = vector.state_vector
    if hasattr(vector, "engine"):
        if not isinstance(vector.hamiltonian.address, BoseFS):
            raise NotImplementedError("state files hold BoseFS addresses only")
        keys, vals = vector.engine.download()
        return vector.hamiltonian.M, vector.hamiltonian.N, keys, vals, vector.style
    if isinstance(vector, DVec):
        items = sorted(vector.storage.items(), key=lambda kv: kv[0].bits)
        if items and not isinstance(items[0][0], BoseFS):
            raise NotImplementedError("state files hold BoseFS addresses only")
        if items:
            M, N = items[0][0].M, items[0][0].N
        else:
            M = N = None
        keys = pack_keys([k.bits for k, _ in items], key_words(M, N) if items else 1)
        integer = isinstance(vector.style, IsStochasticInteger)
        vals = np.array([v for _, v in items], dtype=np.int64 if integer else np.float64)
        return M, N, keys, vals, vector.style
    raise TypeError("save_state takes a DVec, a DeviceDVec or a PMCSimulation")


def state_table(M: int, N: int, keys: np.ndarray, vals: np.ndarray, **metadata):
    """The Arrow table of a state (keys: a word array, (n,) or (n, 2) [lo, hi]; vals: int64
    or float64)."""
    pa, _ = _pa()
    B = M + N - 1
    dt = chunk_dtype(B)
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    nchunks = key_words(M, N)
    if nchunks == 1:
        if B < 64 and len(keys) and int(keys.max()) >> B:
            raise ValueError("key outside the B-bit address space")
        flat = keys.astype(dt, copy=False)
    else:
        keys = keys.reshape(-1, 2)
        if len(keys) and int(keys[:, 1].max()) >> (B - 64):
            raise ValueError("key outside the B-bit address space")
        flat = np.ascontiguousarray(keys[:, ::-1]).reshape(-1)   # chunk 1 = high word first
    chunk_type = pa.from_numpy_dtype(dt)
    key_type = pa.list_(pa.field("item", chunk_type, nullable=False), nchunks)
    key_arr = pa.FixedSizeListArray.from_arrays(pa.array(flat), type=key_type)
    vals = np.ascontiguousarray(vals)
    if vals.dtype not in (np.int64, np.float64):
        raise TypeError("values must be int64 or float64")
    key_field = pa.field("key", key_type, nullable=False, metadata={
        b"ARROW:extension:name": BOSEFS_BS.encode(),
        b"ARROW:extension:metadata": f"{N}.{M}.{B}".encode()})
    val_field = pa.field("value", pa.from_numpy_dtype(vals.dtype), nullable=False)
    meta = {"RIMU_PACKAGE_VERSION": RIMU_PACKAGE_VERSION}
    meta.update({str(k): _julia_string(v) for k, v in metadata.items()})
    schema = pa.schema([key_field, val_field], metadata=meta)
    return pa.Table.from_arrays([key_arr, pa.array(vals)], schema=schema)


def save_state(filename, vector, **metadata):
    """``save_state(filename, vector; metadata...)`` (RimuIO.jl:92-118)."""
    pa, ipc = _pa()
    M, N, keys, vals, _ = _vector_arrays(vector)
    if M is None:
        raise ValueError("an empty host DVec has no address type; save it from a DeviceDVec")
    table = state_table(M, N, keys, vals, **metadata)
    opts = ipc.IpcWriteOptions(compression="zstd", use_threads=True)
    with ipc.new_file(str(filename), table.schema, options=opts) as w:
        w.write_table(table)


def read_state_arrays(filename):
    """(M, N, keys u64 sorted, values, metadata) of a state file."""
    pa, ipc = _pa()
    with pa.memory_map(str(filename), "r") as src:
        table = ipc.open_file(src).read_all()
    if tuple(table.schema.names) != ("key", "value"):
        raise ValueError(f"`{filename}` is not a valid Rimu state file")
    kf = table.schema.field("key")
    fm = kf.metadata or {}
    name = fm.get(b"ARROW:extension:name", b"").decode()
    if name != BOSEFS_BS:
        raise NotImplementedError(f"key type {name or kf.type} is not a one-chunk BoseFS")
    N, M, B = (int(x) for x in fm[b"ARROW:extension:metadata"].decode().split("."))
    nchunks = key_words(M, N)
    if B != N + M - 1 or B > 128 or kf.type.list_size != nchunks:
        raise NotImplementedError("only BoseFS on one or two BitString chunks is supported")
    col = table.column("key").combine_chunks()
    flat = col.flatten().to_numpy(zero_copy_only=False).astype(np.uint64)
    if nchunks == 1:
        keys = flat
        ascending = len(keys) < 2 or bool(np.all(keys[1:] > keys[:-1]))
    else:
        keys = np.ascontiguousarray(flat.reshape(-1, 2)[:, ::-1])   # back to [lo, hi]
        ascending = len(keys) < 2 or bool(np.all(
            (keys[1:, 1] > keys[:-1, 1]) | ((keys[1:, 1] == keys[:-1, 1]) & (keys[1:, 0] > keys[:-1, 0]))))
    vals = table.column("value").combine_chunks().to_numpy(zero_copy_only=False)
    if not ascending:   # files from PDVecs are unsorted
        order = np.argsort(keys, kind="stable") if nchunks == 1 else np.lexsort((keys[:, 0], keys[:, 1]))
        keys, vals = keys[order], vals[order]
    raw = table.schema.metadata or {}
    meta = {k.decode(): _parse_meta(k.decode(), v.decode()) for k, v in raw.items()}
    return M, N, keys, vals, meta


def load_state(filename, style=None):
    """``load_state(filename)`` (RimuIO.jl:130-180): a DVec and the parsed metadata.  The
    default style is IsDynamicSemistochastic for floating values and the default style of
    the value type otherwise."""
    M, N, keys, vals, meta = read_state_arrays(filename)
    if style is None:
        style = IsStochasticInteger() if vals.dtype.kind in "iu" else IsDynamicSemistochastic()
    out = DVec(style=style)
    for k, v in zip(unpack_keys(keys), vals.tolist()):
        out[BoseFS(M=M, N=N, bits=k)] = v
    return out, meta


def save_df(filename, df, compress="auto", metadata=None):
    """``save_df`` (RimuIO.jl:34-56): the report as an Arrow file, zstd above 10,000 rows
    unless ``compress`` says otherwise (``None`` for none, as Julia's ``nothing``), the
    DataFrame's table metadata plus RIMU_PACKAGE_VERSION."""
    pa, ipc = _pa()
    table = pa.Table.from_pandas(df, preserve_index=False)
    meta = dict(metadata) if metadata is not None else {str(k): str(v) for k, v in df.attrs.items()}
    meta["RIMU_PACKAGE_VERSION"] = RIMU_PACKAGE_VERSION
    table = table.replace_schema_metadata({k: str(v) for k, v in meta.items()})
    if compress == "auto":
        compress = "zstd" if len(df) > 10_000 else None
    opts = ipc.IpcWriteOptions(compression=compress)
    with ipc.new_file(str(filename), table.schema, options=opts) as w:
        w.write_table(table)


def load_df(filename, propagate_metadata=True, add_filename=True):
    """``load_df`` (RimuIO.jl:58-79)."""
    pa, ipc = _pa()
    with pa.memory_map(str(filename), "r") as src:
        table = ipc.open_file(src).read_all()
    df = table.to_pandas()
    if propagate_metadata and table.schema.metadata:
        for k, v in table.schema.metadata.items():
            df.attrs[k.decode()] = v.decode()
    if add_filename:
        df.attrs["filename"] = str(filename)
    return df
