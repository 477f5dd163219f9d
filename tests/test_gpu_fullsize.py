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
