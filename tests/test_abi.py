"""CPU checks of the C ABI: the library loads and exports every symbol that
include/rimu_hip.h declares; calls that need a device fail cleanly without one."""
import ctypes
import pathlib
import re

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "rimu_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rimu_[a-z_0-9]+)\s*\(", text)))


def test_header_matches_binding(rimu):
    assert sorted(rimu._native.EXPORTS) == declared_symbols()


def test_library_exports_declared_symbols(rimu):
    lib = rimu._native.lib()
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_nm_shows_exports():
    import subprocess
    so = ROOT / "rimu.jl_amd" / "librimu_hip.so"
    out = subprocess.run(["nm", "-D", "--defined-only", str(so)], capture_output=True, text=True).stdout
    for name in declared_symbols():
        assert re.search(rf"\bT {name}\b", out), name


def test_errors_without_device(rimu):
    if rimu.device_count() > 0:
        pytest.skip("a device is present")
    h = rimu.HubbardReal1D(rimu.near_uniform(4, 4))
    with pytest.raises(rimu.RimuError) as ei:
        rimu._native.Engine(h.descriptor(), rimu.IsStochasticInteger().descriptor(), 1, 10, 10)
    assert ei.value.code == rimu._native.RIMU_ERR_NODEVICE


def test_argument_errors(rimu):
    m = rimu._native.Model()
    m.kind, m.M, m.N = 1, 70, 60  # 129 bits: does not fit two chunks
    s = rimu.IsStochasticInteger().descriptor()
    h = ctypes.c_void_p()
    rc = rimu._native.lib().rimu_engine_create(ctypes.byref(m), ctypes.byref(s), 0, 10, 10, ctypes.byref(h))
    assert rc == rimu._native.RIMU_ERR_ARG
    assert b"two 64-bit chunks" in rimu._native.lib().rimu_last_error()
    m.kind, m.M, m.N = 2, 65, 10  # HubbardMom1D's dispersion table holds 64 modes
    rc = rimu._native.lib().rimu_engine_create(ctypes.byref(m), ctypes.byref(s), 0, 10, 10, ctypes.byref(h))
    assert rc == rimu._native.RIMU_ERR_ARG
    m.kind = 1
    s2 = rimu.IsStochasticInteger().descriptor()
    s2.proj_threshold = 1.0
    m.M, m.N = 5, 5
    rc = rimu._native.lib().rimu_engine_create(ctypes.byref(m), ctypes.byref(s2), 0, 10, 10, ctypes.byref(h))
    assert rc == rimu._native.RIMU_ERR_ARG


def test_product_path_does_not_touch_oracle():
    """The package never imports or loads the oracle."""
    for f in (ROOT / "rimu.jl_amd").rglob("*"):
        if f.suffix in (".py", ".hip", ".hpp", ".cpp", ".h"):
            text = f.read_text()
            assert not re.search(r"import\s+oracle|from\s+oracle|liboracle|oracle\.py", text), f
