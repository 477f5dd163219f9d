"""GPU parity of the step path against the CPU oracle (through the C ABI).

Bars: integer values and addresses bit-exact; floating values bit-exact where the oracle
replays the engine's fold order (list order: diagonal value first, then spawns in
(entry, attempt) order), and within 1e-12 relative of the reference column order.
"""
import zlib

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def random_addresses(M, N, n, rng):
    """n distinct random N-boson M-mode addresses (all of them if the space is small)."""
    import math
    if math.comb(N + M - 1, N) <= 4 * n + 1000:
        basis = O.build_basis(M, N)
        return np.sort(rng.choice(basis, min(n, len(basis)), replace=False))
    out = set()
    while len(out) < n:
        cuts = np.sort(rng.choice(N + M - 1, M - 1, replace=False))
        parts = np.diff(np.concatenate([[-1], cuts, [N + M - 1]])) - 1
        out.add(O.from_onr(parts))
    return np.array(sorted(out), dtype=np.uint64)


def random_vector(M, N, n, rng, kind, scale=5.0):
    keys = random_addresses(M, N, n, rng)
    n = len(keys)
    if kind == "int":
        vals = rng.integers(-int(scale), int(scale) + 1, size=n).astype(np.int64)
        vals[vals == 0] = 1
    else:
        vals = rng.normal(scale=scale, size=n)
    return keys, vals


def oracle_model(kind, M, N, u, t):
    return O.real1d(M, N, u, t) if kind == "real" else O.mom1d(M, N, u, t)


def device_model(rimu, kind, M, N, u, t):
    addr = rimu.near_uniform(N, M)
    return (rimu.HubbardReal1D(addr, u=u, t=t) if kind == "real"
            else rimu.HubbardMom1D(addr, u=u, t=t))


@pytest.mark.parametrize("kind,M,N", [("real", 6, 6), ("real", 10, 10), ("real", 20, 20),
                                      ("mom", 6, 6), ("mom", 16, 16), ("mom", 9, 5)])
def test_hamiltonian_elements(gpu, kind, M, N):
    rng = np.random.default_rng(M * 100 + N)
    h = device_model(gpu, kind, M, N, 4.0, 1.3)
    om = oracle_model(kind, M, N, 4.0, 1.3)
    addrs = random_addresses(M, N, 300, rng)
    diag, nod, _, _ = gpu._native.hamiltonian_eval(h.descriptor(), addrs, np.zeros(len(addrs), np.int64))
    rows, ch = [], []
    for a, n in zip(addrs.tolist(), nod.tolist()):
        assert n == O.num_offdiagonals(om, a)
        rows += [a] * n
        ch += list(range(1, n + 1))
    for a, d in zip(addrs.tolist(), diag.tolist()):
        assert d == O.diagonal_element(om, a)
    _, _, na, ov = gpu._native.hamiltonian_eval(h.descriptor(), np.array(rows, np.uint64), np.array(ch))
    for a, c, b, v in zip(rows, ch, na.tolist(), ov.tolist()):
        ob, oval = O.get_offdiagonal(om, a, c)
        assert b == ob and v == oval, (a, c)


@pytest.mark.parametrize("vtype", ["int", "float"])
@pytest.mark.parametrize("n,nkeys", [(0, 1), (1, 1), (7, 3), (5000, 300), (200_000, 50_000),
                                     (300_000, 7)])
def test_annihilate_bit_exact(gpu, vtype, n, nkeys):
    rng = np.random.default_rng(n + nkeys)
    pool = rng.choice(2 ** 39, nkeys, replace=False).astype(np.uint64)
    keys = pool[rng.integers(0, nkeys, size=n)]
    if vtype == "int":
        vals = rng.integers(-3, 4, size=n).astype(np.int64)
    else:
        vals = rng.normal(size=n) * 10.0 ** rng.integers(-8, 8, size=n)
        # exact cancellations: add the negated value of a few records to the same key
        m = n // 10
        if m:
            idx = rng.integers(0, n, size=m)
            keys = np.concatenate([keys, keys[idx]])
            vals = np.concatenate([vals, -vals[idx]])
    gk, gv = gpu.annihilate(keys, vals)
    ok, ov = O.annihilate(keys, vals)
    assert np.array_equal(gk, ok)
    assert gv.dtype == ov.dtype
    assert np.array_equal(gv.view(np.uint64) if vtype == "float" else gv,
                          ov.view(np.uint64) if vtype == "float" else ov)


STEP_CASES = [
    # model, M, N, u, style, vector length, value scale, shift, dtau[, options]
    ("real", 10, 10, 6.0, "integer", 2000, 6, -2.0, 0.01),
    ("real", 20, 20, 6.0, "integer", 5000, 40, 0.5, 0.001),
    ("mom", 8, 8, 4.0, "integer", 1500, 8, -3.0, 0.005),
    ("mom", 16, 16, 4.0, "integer", 800, 30, -10.0, 0.0005),
    ("real", 10, 10, 6.0, "dynamic", 2000, 4.0, -2.0, 0.01),
    ("real", 20, 20, 6.0, "dynamic", 5000, 60.0, 0.5, 0.001),
    ("mom", 8, 8, 4.0, "dynamic", 1500, 10.0, -3.0, 0.005),
    ("real", 8, 8, 2.0, "deterministic", 1000, 1.0, 0.3, 0.01),
    ("mom", 7, 6, 3.0, "deterministic", 600, 1.0, 0.3, 0.01),
    # IsDynamicSemistochastic(late_compression=false): deposits projected to +-1 when
    # smaller (styles.jl:182-194, spawning.jl:13-31), no ThresholdCompression afterwards
    ("real", 10, 10, 6.0, "dynamic", 2000, 4.0, -2.0, 0.01, {"late_compression": False}),
    ("mom", 8, 8, 4.0, "dynamic", 1500, 3.0, -3.0, 0.005, {"late_compression": False}),
    # non-default spawning thresholds of DynamicSemistochastic (spawning.jl:358-378)
    ("real", 12, 12, 6.0, "dynamic", 3000, 8.0, -1.0, 0.01, {"rel_spawning_threshold": 2.0}),
    ("mom", 8, 8, 4.0, "dynamic", 1500, 10.0, -3.0, 0.005, {"abs_spawning_threshold": 3.0}),
    ("real", 10, 10, 6.0, "dynamic", 2000, 6.0, -2.0, 0.01,
     {"rel_spawning_threshold": 0.5, "abs_spawning_threshold": 5.0, "threshold": 0.7}),
    # boost of apply_operator! (Interfaces/dictvectors.jl:112): more attempts per walker
    ("real", 10, 10, 6.0, "integer", 2000, 6, -2.0, 0.01, {"boost": 2.0}),
    ("mom", 8, 8, 4.0, "dynamic", 1500, 10.0, -3.0, 0.005, {"boost": 2.0}),
    ("real", 12, 12, 6.0, "dynamic", 3000, 3.0, -1.0, 0.01, {"boost": 3.0, "late_compression": False}),
]


def _case_id(c):
    opts = c[9] if len(c) > 9 else {}
    extra = "".join(f"-{k}={v}" for k, v in opts.items())
    return f"{c[0]}{c[1]}x{c[2]}-{c[4]}{extra}"


def _styles(rimu, style, opts):
    """The device style and the oracle style of a case (options as Rimu's keywords)."""
    if style == "integer":
        return rimu.IsStochasticInteger(), O.style("integer")
    if style == "deterministic":
        return rimu.IsDeterministic(), O.style("deterministic")
    kw = {k: opts[k] for k in ("threshold", "rel_spawning_threshold", "abs_spawning_threshold",
                               "late_compression") if k in opts}
    dev = rimu.IsDynamicSemistochastic(**kw)
    th = kw.get("threshold", 1.0)
    late = kw.get("late_compression", True)
    osty = O.style("dynamic", compression_threshold=th, proj_threshold=0.0 if late else th,
                   rel_threshold=kw.get("rel_spawning_threshold", 1.0),
                   abs_threshold=kw.get("abs_spawning_threshold", float("inf")),
                   compression=1 if late else 0)
    return dev, osty


def _engine(rimu, kind, M, N, u, style, seed, opts=None):
    h = device_model(rimu, kind, M, N, u, 1.0)
    st, _ = _styles(rimu, style, opts or {})
    return rimu._native.Engine(h.descriptor(), st.descriptor(), seed, 4_000_000, 1 << 16)


@pytest.mark.parametrize("case", STEP_CASES, ids=_case_id)
def test_step_matches_oracle(gpu, case):
    kind, M, N, u, style, n, scale, shift, dtau = case[:9]
    opts = case[9] if len(case) > 9 else {}
    boost = opts.get("boost", 1.0)
    rng = np.random.default_rng(zlib.crc32(repr(case).encode()))
    keys, vals = random_vector(M, N, n, rng, "int" if style == "integer" else "float", scale)
    seed, step = 0xC0FFEE + n, 37
    eng = _engine(gpu, kind, M, N, u, style, seed, opts)
    eng.upload(keys, vals)
    st = eng.step(step, shift, dtau, boost)
    gk, gv = eng.download()
    eng.close()
    om = oracle_model(kind, M, N, u, 1.0)
    _, osty = _styles(gpu, style, opts)
    ok, ov, ostats = O.step(om, osty, keys, vals, seed=seed, step=step, shift=shift,
                            dtau=dtau, boost=boost, order=1)
    assert np.array_equal(gk, ok)
    if style == "integer":
        assert np.array_equal(gv, ov)
        for name, field in [("attempts", "spawn_attempts"), ("spawns", "spawns"),
                            ("deaths", "deaths"), ("clones", "clones"), ("zombies", "zombies"),
                            ("len", "len"), ("norm", "norm")]:
            assert getattr(st, field) == ostats[name], name
    else:
        assert np.array_equal(gv.view(np.uint64), ov.view(np.uint64))
        assert st.spawn_attempts == ostats["attempts"]
        assert st.exact_steps == ostats["exact_steps"]
        assert st.inexact_steps == ostats["inexact_steps"]
        assert st.len == ostats["len"] and st.len_before == ostats["len_before"]
        assert st.norm == pytest.approx(ostats["norm"], rel=1e-13)
        assert st.spawns == pytest.approx(ostats["spawns"], rel=1e-13)
    # the reference's own column order differs only in floating-point association
    rk, rv, _ = O.step(om, osty, keys, vals, seed=seed, step=step, shift=shift, dtau=dtau,
                       boost=boost, order=0)
    if style == "deterministic":
        assert np.array_equal(rk, gk)
        np.testing.assert_allclose(gv, rv, rtol=1e-12, atol=1e-12 * np.abs(rv).max())


def test_step_sequence_reproducible(gpu):
    """Same seed, same steps: bit-identical vectors (no order dependence on scheduling)."""
    M = N = 12
    rng = np.random.default_rng(5)
    keys, vals = random_vector(M, N, 20000, rng, "float", 20.0)
    outs = []
    for _ in range(2):
        eng = _engine(gpu, "real", M, N, 6.0, "dynamic", 99)
        eng.upload(keys, vals)
        for s in range(1, 6):
            eng.step(s, -1.0, 0.01)
        outs.append(eng.download())
        eng.close()
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1].view(np.uint64), outs[1][1].view(np.uint64))


def test_spawn_buffer_growth_repeats_step(gpu):
    """A step that overflows the spawn buffer is repeated after growing it, with the
    same result as a step that had room from the start."""
    M = N = 10
    rng = np.random.default_rng(11)
    keys, vals = random_vector(M, N, 3000, rng, "int", 50)
    h = device_model(gpu, "real", M, N, 6.0, 1.0)
    sty = gpu.IsStochasticInteger().descriptor()
    small = gpu._native.Engine(h.descriptor(), sty, 5, 100000, 64)
    big = gpu._native.Engine(h.descriptor(), sty, 5, 100000, 1 << 22)
    for e in (small, big):
        e.upload(keys, vals)
        e.step(3, 0.0, 0.01)
    a, b = small.download(), big.download()
    assert small.max_spawns > 64
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    small.close()
    big.close()
