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

FermiFS keys are the same with extension name ``Rimu.FermiFS.BitString`` and B = M
(``arrowtypes.jl:76-81``).  A CompositeFS key (``arrowtypes.jl:107-148``) has extension
name ``Rimu.CompositeFS``, metadata ``"name1:meta1;name2:meta2"`` of its components, and
storage a Struct with fields ``1`` and ``2`` holding each component's chunk tuple (Arrow.jl
stores a tuple of tuples as a named tuple with names 1, 2).

The Arrow conventions follow Arrow.jl (Rimu's dependency, compat "1.5, 2"): an extension
type is the field metadata pair ``ARROW:extension:name`` / ``ARROW:extension:metadata``,
an ``NTuple{N,T}`` is a FixedSizeList<T>[N], and a field is nullable only for a
``Union{Missing,...}`` element type.
"""
from __future__ import annotations

import re

import numpy as np

from .addresses import BoseFS, CompositeFS, FermiFS, key_words, pack_keys, unpack_keys
from .dictvectors import DVec
from .styles import IsDynamicSemistochastic, IsStochasticInteger

RIMU_PACKAGE_VERSION = "0.14.0"   # the Rimu.jl release whose format this mirrors
BOSEFS_BS = "Rimu.BoseFS.BitString"
FERMIFS_BS = "Rimu.FermiFS.BitString"
COMPOSITEFS = "Rimu.CompositeFS"

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
    """(address template, keys (word array), values, style) of a DVec, a DeviceDVec or a
    simulation's state vector; the template is None for an empty host DVec."""
    if hasattr(vector, "state_vector"):
        vector = vector.state_vector
    if hasattr(vector, "engine"):
        keys, vals = vector.engine.download()
        return vector.hamiltonian.address, keys, vals, vector.style
    if isinstance(vector, DVec):
        items = sorted(vector.storage.items(), key=lambda kv: kv[0].bits)
        template = items[0][0] if items else None
        keys = pack_keys([k.bits for k, _ in items], template.key_words if items else 1)
        integer = isinstance(vector.style, IsStochasticInteger)
        vals = np.array([v for _, v in items], dtype=np.int64 if integer else np.float64)
        return template, keys, vals, vector.style
    raise TypeError("save_state takes a DVec, a DeviceDVec or a PMCSimulation")


def _single_spec(addr):
    """(extension name, metadata, chunk dtype, chunks) of a BoseFS or FermiFS on a BitString."""
    B = addr.key_bits
    name = FERMIFS_BS if isinstance(addr, FermiFS) else BOSEFS_BS
    return name, f"{addr.N}.{addr.M}.{B}", chunk_dtype(B), (2 if B > 64 else 1)


def _chunk_array(pa, ints, B):
    """FixedSizeList array of the BitString chunks (chunk 1 = most significant) of B-bit ints."""
    dt = chunk_dtype(B)
    nchunks = 2 if B > 64 else 1
    for x in ints:
        if x >> B:
            raise ValueError("key outside the B-bit address space")
    if nchunks == 1:
        flat = np.array(ints, dtype=np.uint64).astype(dt)
    else:
        flat = np.array([[x >> 64, x & ((1 << 64) - 1)] for x in ints], dtype=np.uint64).reshape(-1)
    ktype = pa.list_(pa.field("item", pa.from_numpy_dtype(dt), nullable=False), nchunks)
    return pa.FixedSizeListArray.from_arrays(pa.array(flat), type=ktype), ktype


def address_state_table(template, keys: np.ndarray, vals: np.ndarray, **metadata):
    """The Arrow table of a state whose addresses have the type of ``template`` (a BoseFS,
    FermiFS or two-component CompositeFS); keys are a word array of those addresses."""
    pa, _ = _pa()
    if isinstance(template, BoseFS) and not isinstance(template, CompositeFS):
        return state_table(template.M, template.N, keys, vals, **metadata)
    vals = np.ascontiguousarray(vals)
    if vals.dtype not in (np.int64, np.float64):
        raise TypeError("values must be int64 or float64")
    ints = unpack_keys(np.ascontiguousarray(keys, dtype=np.uint64))
    if isinstance(template, FermiFS):
        name, meta_s, _, _ = _single_spec(template)
        key_arr, ktype = _chunk_array(pa, ints, template.key_bits)
    elif isinstance(template, CompositeFS):
        a, b = template.components
        wb = b.key_bits
        arr_a, ta = _chunk_array(pa, [x >> wb for x in ints], a.key_bits)
        arr_b, tb = _chunk_array(pa, [x & ((1 << wb) - 1) for x in ints], b.key_bits)
        fields = [pa.field("1", ta, nullable=False), pa.field("2", tb, nullable=False)]
        key_arr = pa.StructArray.from_arrays([arr_a, arr_b], fields=fields)
        ktype = key_arr.type
        name = COMPOSITEFS
        meta_s = ";".join(f"{n}:{m}" for n, m, _, _ in (_single_spec(a), _single_spec(b)))
    else:
        raise TypeError(f"unsupported address type {type(template).__name__}")
    key_field = pa.field("key", ktype, nullable=False, metadata={
        b"ARROW:extension:name": name.encode(), b"ARROW:extension:metadata": meta_s.encode()})
    val_field = pa.field("value", pa.from_numpy_dtype(vals.dtype), nullable=False)
    meta = {"RIMU_PACKAGE_VERSION": RIMU_PACKAGE_VERSION}
    meta.update({str(k): _julia_string(v) for k, v in metadata.items()})
    schema = pa.schema([key_field, val_field], metadata=meta)
    return pa.Table.from_arrays([key_arr, pa.array(vals)], schema=schema)


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
    template, keys, vals, _ = _vector_arrays(vector)
    if template is None:
        raise ValueError("an empty host DVec has no address type; save it from a DeviceDVec")
    table = address_state_table(template, keys, vals, **metadata)
    opts = ipc.IpcWriteOptions(compression="zstd", use_threads=True)
    with ipc.new_file(str(filename), table.schema, options=opts) as w:
        w.write_table(table)


def _read_table(filename):
    pa, ipc = _pa()
    with pa.memory_map(str(filename), "r") as src:
        table = ipc.open_file(src).read_all()
    if tuple(table.schema.names) != ("key", "value"):
        raise ValueError(f"`{filename}` is not a valid Rimu state file")
    return table


def _single_from_meta(name, meta_s):
    N, M, B = (int(x) for x in meta_s.split("."))
    if name == BOSEFS_BS and B == N + M - 1 and B <= 128:
        return BoseFS(M=M, N=N, bits=0)
    if name == FERMIFS_BS and B == M and M <= 128:
        return FermiFS(M=M, N=N, bits=0)
    raise NotImplementedError(f"key type {name} {meta_s} is not supported")


def _chunk_ints(col, B):
    nchunks = 2 if B > 64 else 1
    if col.type.list_size != nchunks:
        raise NotImplementedError("unexpected number of BitString chunks")
    flat = col.flatten().to_numpy(zero_copy_only=False).astype(np.uint64)
    if nchunks == 1:
        return [int(x) for x in flat.tolist()]
    return [(int(hi) << 64) | int(lo) for hi, lo in flat.reshape(-1, 2).tolist()]


def read_state(filename):
    """(address template, keys as a sorted word array, values, metadata) of a state file of
    BoseFS, FermiFS or two-component CompositeFS addresses."""
    table = _read_table(filename)
    kf = table.schema.field("key")
    fm = kf.metadata or {}
    name = fm.get(b"ARROW:extension:name", b"").decode()
    meta_s = fm.get(b"ARROW:extension:metadata", b"").decode()
    col = table.column("key").combine_chunks()
    if name in (BOSEFS_BS, FERMIFS_BS):
        template = _single_from_meta(name, meta_s)
        ints = _chunk_ints(col, template.key_bits)
    elif name == COMPOSITEFS:
        parts = [p.split(":", 1) for p in meta_s.split(";")]
        if len(parts) != 2:
            raise NotImplementedError("CompositeFS state files of two components are supported")
        a, b = (_single_from_meta(n, m) for n, m in parts)
        if a.M != b.M:
            raise ValueError("components with different numbers of modes")
        ia = _chunk_ints(col.field(0), a.key_bits)
        ib = _chunk_ints(col.field(1), b.key_bits)
        template = CompositeFS(a, b)
        ints = [(x << b.key_bits) | y for x, y in zip(ia, ib)]
    else:
        raise NotImplementedError(f"key type {name or kf.type} is not supported")
    vals = table.column("value").combine_chunks().to_numpy(zero_copy_only=False)
    order = sorted(range(len(ints)), key=ints.__getitem__)   # files from PDVecs are unsorted
    keys = pack_keys([ints[i] for i in order], template.key_words)
    vals = vals[np.array(order, dtype=np.int64)] if len(order) else vals
    raw = table.schema.metadata or {}
    meta = {k.decode(): _parse_meta(k.decode(), v.decode()) for k, v in raw.items()}
    return template, keys, vals, meta


def read_state_arrays(filename):
    """(M, N, keys u64 sorted, values, metadata) of a BoseFS state file."""
    pa, ipc = _pa()
    table = _read_table(filename)
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
    template, keys, vals, meta = read_state(filename)
    if style is None:
        style = IsStochasticInteger() if vals.dtype.kind in "iu" else IsDynamicSemistochastic()
    out = DVec(style=style)
    for k, v in zip(unpack_keys(keys), vals.tolist()):
        out[template.with_bits(k)] = v
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
