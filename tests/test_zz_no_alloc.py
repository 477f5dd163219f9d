"""rimu_annihilate_noalloc against rimu_annihilate through the C ABI: the same output on
two calls in a row (the second with a smaller list, so stale scratch contents of the first
call would show), for int64 and double values, and the size limit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _list(rng, n, nkeys, vtype):
    pool = rng.choice(2 ** 39, nkeys, replace=False).astype(np.uint64)
    keys = pool[rng.integers(0, nkeys, size=n)]
    if vtype == "int":
        vals = rng.integers(-3, 4, size=n).astype(np.int64)
    else:
        vals = rng.normal(size=n) * 10.0 ** rng.integers(-8, 8, size=n)
        m = n // 10
        idx = rng.integers(0, n, size=m)
        keys = np.concatenate([keys, keys[idx]])
        vals = np.concatenate([vals, -vals[idx]])
    return keys, vals


def _same(a, b, vtype):
    assert np.array_equal(a[0], b[0])
    if vtype == "float":
        assert np.array_equal(a[1].view(np.uint64), b[1].view(np.uint64))
    else:
        assert np.array_equal(a[1], b[1])


@pytest.mark.parametrize("vtype", ["int", "float"])
def test_noalloc_matches_annihilate_twice(gpu, vtype):
    rng = np.random.default_rng(2024 if vtype == "int" else 2025)
    for n, nkeys in [(900_000, 200_000), (30_000, 40)]:
        keys, vals = _list(rng, n, nkeys, vtype)
        print(f"noalloc {vtype} n={len(keys)} nkeys={nkeys}", flush=True)
        ref = gpu._native.annihilate(keys, vals)
        got = gpu._native.annihilate(keys, vals, noalloc=True)
        _same(got, ref, vtype)


def test_noalloc_size_limit(gpu):
    n = (1 << 20) + 1
    keys = np.arange(n, dtype=np.uint64)
    vals = np.ones(n, dtype=np.int64)
    with pytest.raises(gpu.RimuError) as ei:
        gpu._native.annihilate(keys, vals, noalloc=True)
    assert ei.value.code == gpu._native.RIMU_ERR_CAPACITY
    # the largest list it takes, every key distinct
    keys = np.arange(n - 1, dtype=np.uint64)[::-1].copy() * np.uint64(3)
    vals = np.arange(1, n, dtype=np.int64)
    got = gpu._native.annihilate(keys, vals, noalloc=True)
    assert np.array_equal(got[0], np.sort(keys))
    assert np.array_equal(got[1], vals[::-1])
