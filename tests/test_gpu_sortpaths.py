"""The step's sort and merge paths through the C ABI: the bucket merge (sort on the top bits
of the stored keys, each bucket sorted and folded in LDS) against the merge path after a sort
on every bit, the fallback of an overflowing bucket, and the bounded look-back waits giving up
when forced from the host (the step is then repeated with the per-pass sort)."""
import numpy as np
import pytest

import oracle as O
from test_gpu_parity import device_model, oracle_model, random_vector

pytestmark = pytest.mark.gpu


def _engine(rimu, kind, M, N, u, style, seed, mode, rule=0, thr=0.0):
    h = device_model(rimu, kind, M, N, u, 1.0)
    st = {"integer": rimu.IsStochasticInteger(), "dynamic": rimu.IsDynamicSemistochastic()}[style]
    e = rimu._native.Engine(h.descriptor(), st.descriptor(), seed, 8_000_000, 1 << 20)
    e.set_radix_mode(mode)
    if rule:
        e.set_initiator(rule, thr)
    e.set_timing(True)
    return e


def _kernels(e):
    return set(e.kernel_times().keys())


@pytest.mark.parametrize("kind,M,N,style,n,scale", [
    ("real", 20, 20, "dynamic", 200_000, 60.0),    # 39-bit keys: two top-bit passes
    ("real", 20, 20, "integer", 100_000, 200),
    ("mom", 16, 16, "dynamic", 60_000, 40.0),      # 31-bit keys
    ("real", 14, 12, "dynamic", 50_000, 10.0),     # 25-bit keys: one pass
])
def test_bucket_merge_equals_merge_path(gpu, kind, M, N, style, n, scale):
    """Five steps on the bucket path and on the merge path from the same vector and seed:
    bit-identical vectors and statistics."""
    rng = np.random.default_rng(n + M)
    keys, vals = random_vector(M, N, n, rng, "int" if style == "integer" else "float", scale)
    print(f"bucket vs merge path {kind} {M}x{N} {style} n={len(keys)}", flush=True)
    out = []
    for mode in (2, 1):
        e = _engine(gpu, kind, M, N, 4.0, style, 71, mode)
        e.upload(keys, vals)
        stats = [e.step(s, -2.0, 1e-3) for s in range(1, 6)]
        names = _kernels(e)
        out.append((e.download(), [(st.len, st.norm, st.spawns, st.len_before) for st in stats], names))
        e.close()
    (k2, v2), s2, names2 = out[0]
    (k1, v1), s1, names1 = out[1]
    assert "bucket_merge" in names2 and "bucket_fallback" not in names2
    assert "bucket_merge" not in names1
    assert np.array_equal(k2, k1)
    assert np.array_equal(v2.view(np.uint64), v1.view(np.uint64))
    # norms sum in another order (per bucket against per merge tile)
    for a, b in zip(s2, s1):
        assert a[0] == b[0] and a[3] == b[3]
        assert a[1] == pytest.approx(b[1], rel=1e-13) and a[2] == pytest.approx(b[2], rel=1e-13)


def test_bucket_overflow_falls_back(gpu):
    """Ten entries of 2e5 walkers spawn about 1e6 records onto a few hundred addresses: their
    buckets outgrow the LDS tiles, the step falls back to the merge path, and the result
    equals the oracle's."""
    M = N = 20
    rng = np.random.default_rng(4)
    keys, _ = random_vector(M, N, 10, rng, "int", 1)
    vals = rng.integers(150_000, 250_000, size=len(keys)) * rng.choice([-1, 1], size=len(keys))
    e = _engine(gpu, "real", M, N, 6.0, "integer", 123, 2)
    e.upload(keys, vals)
    st = e.step(9, -1.0, 0.01)
    gk, gv = e.download()
    names = _kernels(e)
    e.close()
    assert "bucket_merge" in names and "bucket_fallback" in names
    ok, ov, ost = O.step(oracle_model("real", M, N, 6.0, 1.0), O.style("integer"), keys, vals, seed=123,
                         step=9, shift=-1.0, dtau=0.01, order=1)
    assert np.array_equal(gk, ok) and np.array_equal(gv, ov)
    assert st.spawns == ost["spawns"] and st.norm == ost["norm"] and st.len == ost["len"]


@pytest.mark.parametrize("mode", [2, 1])
def test_lookback_give_up_repeats_step(gpu, mode):
    """The first tile of every radix pass withholds its counts (a host-forced test): every later
    tile's bounded wait gives up, every kernel exits, and the engine repeats the step with the
    per-pass sort; the result equals a fresh engine's step."""
    M = N = 20
    rng = np.random.default_rng(8)
    keys, vals = random_vector(M, N, 40_000, rng, "float", 30.0)
    e = _engine(gpu, "real", M, N, 6.0, "dynamic", 5, mode)
    e.upload(keys, vals)
    e.set_lookback(1000, withhold_first=True)
    st = e.step(3, -1.0, 1e-3)
    assert "lookback_fallback" in _kernels(e)
    got = e.download()  # its device sort gives up too, and sorts again with the per-pass sort
    e.set_lookback(1 << 25, withhold_first=False)
    assert np.array_equal(e.download()[0], got[0])
    e.close()
    f = _engine(gpu, "real", M, N, 6.0, "dynamic", 5, mode)
    f.upload(keys, vals)
    sf = f.step(3, -1.0, 1e-3)
    want = f.download()
    assert "lookback_fallback" not in _kernels(f)
    f.close()
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1].view(np.uint64), want[1].view(np.uint64))
    assert st.len == sf.len and st.norm == pytest.approx(sf.norm, rel=1e-13)


@pytest.mark.parametrize("rule", [1, 2, 3])
@pytest.mark.parametrize("kind,M,N,style,n,scale,thr", [
    ("real", 20, 20, "dynamic", 150_000, 60.0, 20.0),
    ("real", 20, 20, "integer", 80_000, 200, 100.0),
    ("mom", 16, 16, "dynamic", 50_000, 40.0, 10.0),
])
def test_initiator_bucket_merge_equals_merge_path(gpu, kind, M, N, style, n, scale, thr, rule):
    """Initiator rules (Initiator, SimpleInitiator, CoherentInitiator) on the bucket path
    (three components folded per address in LDS) and on the merge path: five steps from the
    same vector and seed give bit-identical vectors and statistics."""
    rng = np.random.default_rng(n + M + rule)
    keys, vals = random_vector(M, N, n, rng, "int" if style == "integer" else "float", scale)
    print(f"initiator bucket vs merge path {kind} {M}x{N} {style} rule {rule} n={len(keys)}", flush=True)
    out = []
    for mode in (2, 1):
        e = _engine(gpu, kind, M, N, 4.0, style, 73, mode, rule, thr)
        e.upload(keys, vals)
        stats = [e.step(s, -2.0, 1e-3) for s in range(1, 6)]
        names = _kernels(e)
        out.append((e.download(), [(st.len, st.norm, st.spawns, st.len_before) for st in stats], names))
        e.close()
    (k2, v2), s2, names2 = out[0]
    (k1, v1), s1, names1 = out[1]
    assert "bucket_merge" in names2 and "bucket_fallback" not in names2
    assert "bucket_merge" not in names1
    assert np.array_equal(k2, k1)
    assert np.array_equal(v2.view(np.uint64), v1.view(np.uint64))
    for a, b in zip(s2, s1):
        assert a[0] == b[0] and a[3] == b[3]
        assert a[1] == pytest.approx(b[1], rel=1e-13) and a[2] == pytest.approx(b[2], rel=1e-13)
