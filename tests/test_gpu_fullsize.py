"""Checks at the workloads' full sizes, through properties that do not depend on size:
the 1e8-walker HubbardReal1D M = N = 20 state (sorted, unique, nonzero, valid addresses;
len, norm and the attempt statistics against an independent computation over every
entry; the same seed gives the same state), and annihilation of 5e7 integer records
against numpy's grouped sums (integer sums are exact in any order)."""
import hashlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def grow_headline(rimu, steps):
    """The headline run from Rimu's default start (near_uniform, population 10, shift 0),
    DoubleLogUpdate to 1e8 walkers, dtau 1e-3, seed 17."""
    M = N = 20
    H = rimu.HubbardReal1D(rimu.near_uniform(N, M), u=6.0, t=1.0)
    target = 1e8
    eng = rimu._native.Engine(H.descriptor(), rimu.IsDynamicSemistochastic().descriptor(), 17,
                              int(2 * target + 101), int(4 * target + 4096))
    start = H.starting_address()
    eng.upload(np.array([start.bits], dtype=np.uint64), np.array([10.0]))
    strat = rimu.DoubleLogUpdate(target_walkers=target)
    sp = rimu.strategies.DefaultShiftParameters(H.diagonal_element(start), 10.0, 1e-3)
    st = None
    for step in range(1, steps + 1):
        st = eng.step(step, sp.shift, 1e-3)
        strat.update(sp, st.norm)
    return eng, sp, st


def digest(keys, vals):
    h = hashlib.sha256()
    h.update(keys.tobytes())
    h.update(vals.tobytes())
    return h.hexdigest()


def test_headline_state_properties_and_attempt_counts(gpu):
    M = N = 20
    B = M + N - 1
    eng, sp, st = grow_headline(gpu, 400)
    keys, vals = eng.download()
    assert len(keys) > 1e7 and st.norm > 5e7
    assert len(keys) == st.len
    assert np.all(keys[1:] > keys[:-1])                       # sorted, unique
    assert np.all(vals != 0)
    assert np.all(np.bitwise_count(keys) == N) and np.all(keys < np.uint64(1) << np.uint64(B))
    assert np.abs(vals).sum() == pytest.approx(st.norm, rel=1e-12)
    # the next step's attempts: exact columns (|v| >= number of off-diagonals) spawn to
    # every neighbour, the others make max(floor(|v|), 1) attempts (styles.jl:182-214)
    n_od = 2 * np.bitwise_count(keys & ~(keys << np.uint64(1))).astype(np.int64)
    av = np.abs(vals)
    exact = av >= n_od
    attempts = np.where(exact, n_od, np.maximum(np.floor(av), 1).astype(np.int64))
    st2 = eng.step(401, sp.shift, 1e-3)
    assert st2.spawn_attempts == float(attempts.sum())
    assert st2.exact_steps == float(exact.sum())
    assert st2.inexact_steps == float((~exact).sum())
    k2, v2 = eng.download()
    assert len(k2) == st2.len and np.all(k2[1:] > k2[:-1]) and np.all(v2 != 0)
    assert np.abs(v2).sum() == pytest.approx(st2.norm, rel=1e-12)
    # ThresholdCompression leaves no magnitude below 1
    assert np.all(np.abs(v2) >= 1.0)
    d1 = digest(k2, v2)
    eng.close()
    # same seed, same steps: the same state, bit for bit
    eng, sp, st = grow_headline(gpu, 401)
    k3, v3 = eng.download()
    eng.close()
    assert digest(k3, v3) == d1


def test_annihilate_fifty_million_integer_records(gpu):
    rng = np.random.default_rng(2024)
    n = 50_000_000
    pool = np.unique(rng.integers(1, 1 << 39, size=12_000_000, dtype=np.uint64))
    keys = pool[rng.integers(0, len(pool), size=n)]
    vals = rng.integers(-3, 4, size=n).astype(np.int64)
    ok, ov = gpu.annihilate(keys, vals)
    order = np.argsort(keys, kind="stable")
    ks = keys[order]
    vs = vals[order]
    starts = np.flatnonzero(np.concatenate([[True], ks[1:] != ks[:-1]]))
    sums = np.add.reduceat(vs, starts)
    uk = ks[starts]
    keep = sums != 0
    assert np.array_equal(ok, uk[keep])
    assert np.array_equal(ov, sums[keep])


def random_bose_keys(M, N, n, rng, chunk=250_000):
    """About n distinct random BoseFS{N,M} addresses of N + M - 1 in 65..128 bits as (lo, hi)
    word rows, ascending: any bit string with N ones and M - 1 zeros is one."""
    B = N + M - 1
    shifts = np.arange(64, dtype=np.uint64)
    rows = []
    for a in range(0, n, chunk):
        m = min(chunk, n - a)
        ones = np.argsort(rng.random((m, B)), axis=1)[:, :N]
        bits = np.zeros((m, 128), dtype=np.uint64)
        np.put_along_axis(bits, ones, np.uint64(1), axis=1)
        lo = (bits[:, :64] << shifts).sum(axis=1, dtype=np.uint64)
        hi = (bits[:, 64:] << shifts).sum(axis=1, dtype=np.uint64)
        rows.append(np.stack([lo, hi], axis=1))
    keys = np.unique(np.concatenate(rows), axis=0)      # rows in (lo, hi) lexicographic order
    return keys[np.lexsort((keys[:, 0], keys[:, 1]))]


def _step_both_modes(rimu, make_engine, keys, vals, steps, shift, dtau, sort_bits=None, modes=(0, 1)):
    """The same state stepped by one engine per sort mode; returns every result."""
    out = []
    for mode in modes:
        eng = make_engine()
        eng.set_radix_mode(mode)
        if sort_bits is not None:
            eng.set_sort_bits(sort_bits)
        eng.upload(keys, vals)
        stats = [eng.step(s, shift, dtau) for s in range(1, steps + 1)]
        k, v = eng.download()
        eng.close()
        out.append((k, v, stats))
    return out


def _assert_same(a, b, norm_rel=0.0):
    """Same vectors bit for bit and the same statistics; norm_rel > 0 allows the norm's sum
    to differ in association (the bucket merge sums per bucket, the merge path per tile)."""
    (k0, v0, s0), (k1, v1, s1) = a, b
    assert k0.shape == k1.shape and np.array_equal(k0, k1)
    assert np.array_equal(v0.view(np.uint64), v1.view(np.uint64))  # bit for bit
    for x, y in zip(s0, s1):
        for f in ("spawn_attempts", "spawns", "deaths", "clones", "zombies", "len", "len_before"):
            assert getattr(x, f) == getattr(y, f), f
        assert x.norm == pytest.approx(y.norm, rel=norm_rel, abs=0)


def test_radix_modes_agree_on_headline_state(gpu):
    """The bucket merge (sort on the top 16 bits, 65536 buckets sorted and folded in LDS), the
    look-back radix sort (one histogram pass, tiles ordered by start) and the per-pass count
    sort give the same steps, bit for bit, on the 1e8-walker state (about 7e7 spawn records,
    17000 radix tiles per pass)."""
    eng, sp, st = grow_headline(gpu, 300)
    keys, vals = eng.download()
    eng.close()
    assert len(keys) > 1e7
    M = N = 20
    H = gpu.HubbardReal1D(gpu.near_uniform(N, M), u=6.0, t=1.0)

    def make():
        return gpu._native.Engine(H.descriptor(), gpu.IsDynamicSemistochastic().descriptor(), 91,
                                  int(2e8 + 101), int(4e8 + 4096))

    a, b, c = _step_both_modes(gpu, make, keys, vals, 2, sp.shift, 1e-3, modes=(0, 1, 2))
    _assert_same(a, b)
    _assert_same(b, c, norm_rel=1e-13)


@pytest.mark.parametrize("sort_bits", [40, 128])
def test_radix_modes_agree_on_two_word_keys(gpu, sort_bits):
    """Both radix modes on 128-bit keys (HubbardReal1D M = N = 50, 3e6 entries): the top-bit
    sort with its run ordering (40) and the full 16-pass sort (128)."""
    rng = np.random.default_rng(5050 + sort_bits)
    keys = random_bose_keys(50, 50, 3_000_000, rng)
    vals = rng.normal(scale=6.0, size=len(keys))
    H = gpu.HubbardReal1D(gpu.near_uniform(50, 50), u=4.0, t=1.0)

    def make():
        return gpu._native.Engine(H.descriptor(), gpu.IsDynamicSemistochastic().descriptor(), 77,
                                  40_000_000, 80_000_000)

    a, b = _step_both_modes(gpu, make, keys, vals, 2, -2.0, 2e-3, sort_bits)
    _assert_same(a, b)
    assert len(a[0]) > 2_900_000


def test_radix_modes_agree_while_growing(gpu):
    """Both radix modes from Rimu's default start (one address, population 10) through 300
    steps of growth to 2e5 walkers: every step's statistics and the final state agree bit
    for bit.  The look-back status array grows many times and is reused by sorts of other
    sizes (words of earlier sorts carry older epochs)."""
    M = N = 20
    H = gpu.HubbardReal1D(gpu.near_uniform(N, M), u=6.0, t=1.0)
    start = H.starting_address()
    runs = []
    shifts = []
    for mode in (0, 1, 2):
        eng = gpu._native.Engine(H.descriptor(), gpu.IsDynamicSemistochastic().descriptor(), 23, 400_101, 800_000)
        eng.set_radix_mode(mode)
        eng.upload(np.array([start.bits], dtype=np.uint64), np.array([10.0]))
        strat = gpu.DoubleLogUpdate(target_walkers=2e5)
        sp = gpu.strategies.DefaultShiftParameters(H.diagonal_element(start), 10.0, 1e-3)
        stats = []
        for step in range(1, 301):
            # the bucket merge sums the norm in another association: it is given the shifts
            # of the merge path, so that its vectors can be compared bit for bit
            shift = shifts[step - 1] if mode == 2 else sp.shift
            if mode == 1:
                shifts.append(sp.shift)
            st = eng.step(step, shift, 1e-3)
            strat.update(sp, st.norm)
            stats.append((st.spawn_attempts, st.spawns, st.len, st.norm))
        k, v = eng.download()
        eng.close()
        runs.append((k, v, stats))
    (k0, v0, s0), (k1, v1, s1), (k2, v2, s2) = runs
    assert s0 == s1
    assert np.array_equal(k0, k1) and np.array_equal(v0.view(np.uint64), v1.view(np.uint64))
    assert np.array_equal(k0, k2) and np.array_equal(v0.view(np.uint64), v2.view(np.uint64))
    # the shift follows the norms, whose sums differ in association on the bucket path
    for a, b in zip(s1, s2):
        assert a[:3] == b[:3] and a[3] == pytest.approx(b[3], rel=1e-13)
    assert s0[-1][3] > 1e5


def test_radix_mode_rejects_unknown_values(gpu):
    H = gpu.HubbardReal1D(gpu.near_uniform(4, 4), u=1.0, t=1.0)
    eng = gpu._native.Engine(H.descriptor(), gpu.IsDynamicSemistochastic().descriptor(), 1, 1000, 1000)
    try:
        for bad in (-1, 3, 7):
            with pytest.raises(gpu.RimuError):
                eng.set_radix_mode(bad)
        eng.set_radix_mode(0)
        eng.set_radix_mode(1)
        eng.set_radix_mode(2)
    finally:
        eng.close()
