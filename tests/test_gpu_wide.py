"""GPU parity of two-chunk addresses (64 < B = N + M - 1 <= 128) against the oracle: the
engine keeps them as 128-bit keys, the C ABI passes two 64-bit words per address (low word
first).

Bars as for one chunk: addresses and integer values bit-exact; floating values bit-exact
against the oracle's list order; matrix elements bit-exact; overlaps within 1e-12.
"""
import zlib

import numpy as np
import pytest

import oracle as O
from test_gpu_parity import device_model, oracle_model

pytestmark = pytest.mark.gpu

STYLES = {"integer": "IsStochasticInteger", "dynamic": "IsDynamicSemistochastic",
          "deterministic": "IsDeterministic"}


def wide_addresses(M, N, n, rng):
    """n distinct random addresses as Python ints, ascending."""
    B = N + M - 1
    out = set()
    while len(out) < n:
        cuts = np.sort(rng.choice(B, M - 1, replace=False))
        out.add(O.from_onr(np.diff(np.concatenate([[-1], cuts, [B]])) - 1))
    return sorted(out)


def wide_vector(M, N, n, rng, vtype, scale):
    ints = wide_addresses(M, N, n, rng)
    keys = O.pack_keys(ints, 2)
    if vtype == "int":
        vals = rng.integers(-int(scale), int(scale) + 1, size=n).astype(np.int64)
        vals[vals == 0] = 1
    else:
        vals = rng.normal(scale=scale, size=n)
    return keys, vals


def _engine(rimu, kind, M, N, u, style, seed):
    h = device_model(rimu, kind, M, N, u, 1.0)
    st = getattr(rimu, STYLES[style])()
    return rimu._native.Engine(h.descriptor(), st.descriptor(), seed, 4_000_000, 1 << 16)


def _compare(style, gk, gv, st, ok, ov, ostats):
    assert gk.shape == ok.shape and np.array_equal(gk, ok)
    if style == "integer":
        assert np.array_equal(gv, ov)
        for name, field in [("attempts", "spawn_attempts"), ("spawns", "spawns"),
                            ("deaths", "deaths"), ("clones", "clones"), ("zombies", "zombies"),
                            ("len", "len"), ("len_before", "len_before"), ("norm", "norm")]:
            assert getattr(st, field) == ostats[name], name
    else:
        assert np.array_equal(gv == 0, ov == 0)
        nz = ov != 0
        assert np.array_equal(gv[nz].view(np.uint64), ov[nz].view(np.uint64))
        assert st.spawn_attempts == ostats["attempts"]
        assert st.exact_steps == ostats["exact_steps"]
        assert st.inexact_steps == ostats["inexact_steps"]
        assert st.len == ostats["len"] and st.len_before == ostats["len_before"]
        # norm and spawns are sums over up to 1e6 terms, reduced on the device in another
        # order than the oracle's loop: n * eps bounds their relative difference
        assert st.norm == pytest.approx(ostats["norm"], rel=1e-11)
        assert st.spawns == pytest.approx(ostats["spawns"], rel=1e-11)


@pytest.mark.parametrize("kind,M,N", [("real", 50, 50), ("real", 100, 28), ("real", 64, 64),
                                      ("real", 65, 64), ("mom", 40, 30), ("mom", 64, 64)])
def test_hamiltonian_elements_two_chunks(gpu, kind, M, N):
    rng = np.random.default_rng(M * 1000 + N)
    h = device_model(gpu, kind, M, N, 4.0, 1.3)
    om = oracle_model(kind, M, N, 4.0, 1.3)
    ints = wide_addresses(M, N, 200, rng)
    addrs = O.pack_keys(ints, 2)
    diag, nod, _, _ = gpu._native.hamiltonian_eval(h.descriptor(), addrs, np.zeros(len(ints), np.int64))
    rows, ch = [], []
    for a, d, n in zip(ints, diag.tolist(), nod.tolist()):
        assert n == O.num_offdiagonals(om, a)
        assert d == O.diagonal_element(om, a)
        # a spread of channels per address, the first and last included
        cs = sorted({1, n} | set(rng.integers(1, n + 1, size=min(n, 12)).tolist()))
        rows += [a] * len(cs)
        ch += cs
    _, _, na, ov = gpu._native.hamiltonian_eval(h.descriptor(), O.pack_keys(rows, 2), np.array(ch))
    got = O.unpack_keys(na)
    for a, c, b, v in zip(rows, ch, got, ov.tolist()):
        ob, oval = O.get_offdiagonal(om, a, c)
        assert b == ob and v == oval, (a, c)


CASES = [
    # model, M, N, u, style, vector length, value scale, shift, dtau
    ("real", 50, 50, 6.0, "integer", 3000, 8, -2.0, 1e-3),
    ("real", 50, 50, 6.0, "dynamic", 3000, 20.0, -2.0, 1e-3),
    ("real", 50, 50, 6.0, "deterministic", 20, 1.0, 0.3, 1e-2),
    ("real", 64, 64, 6.0, "integer", 2000, 8, -2.0, 1e-3),          # B = 127
    ("real", 65, 64, 6.0, "integer", 2000, 8, -2.0, 1e-3),          # B = 128
    ("real", 65, 64, 6.0, "dynamic", 2000, 20.0, -2.0, 1e-3),
    ("real", 100, 28, 2.0, "dynamic", 2000, 5.0, 0.0, 1e-2),
    ("mom", 40, 30, 4.0, "integer", 1500, 8, -3.0, 5e-3),
    ("mom", 40, 30, 4.0, "dynamic", 1500, 10.0, -3.0, 5e-3),
    ("mom", 64, 64, 4.0, "dynamic", 500, 10.0, -3.0, 1e-3),         # B = 127
    ("mom", 30, 40, 3.0, "deterministic", 40, 1.0, 0.3, 1e-2),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}{c[1]}x{c[2]}-{c[4]}")
def test_two_chunk_steps_match_oracle(gpu, case):
    """Steps in a row, each from the engine's own previous vector (deterministic vectors
    grow by the number of off-diagonals per step, so those take fewer steps)."""
    kind, M, N, u, style, n, scale, shift, dtau = case
    rng = np.random.default_rng(zlib.crc32(repr(case).encode()))
    keys, vals = wide_vector(M, N, n, rng, "int" if style == "integer" else "float", scale)
    seed = 0x5EED + n
    eng = _engine(gpu, kind, M, N, u, style, seed)
    assert eng.kw == 2
    eng.upload(keys, vals)
    om = oracle_model(kind, M, N, u, 1.0)
    ck, cv = keys, vals
    steps = (3, 4, 5) if style != "deterministic" else (3, 4) if kind == "real" else (3,)
    for step in steps:
        st = eng.step(step, shift, dtau)
        gk, gv = eng.download()
        ok, ov, ostats = O.step(om, O.style(style), ck, cv, seed=seed, step=step, shift=shift,
                                dtau=dtau, order=1)
        _compare(style, gk, gv, st, ok, ov, ostats)
        ck, cv = gk, gv
    eng.close()
    # the high words are in use: addresses above 2^64 are present
    assert np.any(ck[:, 1] != 0)


@pytest.mark.parametrize("rule", ["initiator", "simple", "coherent"])
@pytest.mark.parametrize("case", [("real", 50, 50, "integer", 2.0), ("real", 50, 50, "dynamic", 10.0),
                                  ("real", 64, 64, "integer", 3.0), ("mom", 40, 30, "dynamic", 5.0)],
                         ids=lambda c: f"{c[0]}{c[1]}x{c[2]}-{c[3]}")
def test_two_chunk_initiator_steps_match_oracle(gpu, case, rule):
    """The class bit sits at bit 127 of the 128-bit key (B <= 127)."""
    kind, M, N, style, thr = case
    rng = np.random.default_rng(zlib.crc32(repr((case, rule)).encode()))
    keys, vals = wide_vector(M, N, 2000, rng, "int" if style == "integer" else "float",
                             6 if style == "integer" else 20.0)
    seed = 0xA11 + M
    eng = _engine(gpu, kind, M, N, 4.0, style, seed)
    eng.set_initiator({"initiator": 1, "simple": 2, "coherent": 3}[rule], thr)
    eng.upload(keys, vals)
    om = oracle_model(kind, M, N, 4.0, 1.0)
    ck, cv = keys, vals
    for step in (7, 8):
        st = eng.step(step, -2.0, 2e-3)
        gk, gv = eng.download()
        ok, ov, ostats = O.step_initiator(om, O.style(style), ck, cv, rule=rule, threshold=thr,
                                          seed=seed, step=step, shift=-2.0, dtau=2e-3, order=1)
        _compare(style, gk, gv, st, ok, ov, ostats)
        ck, cv = gk, gv
    eng.close()


def test_two_chunk_initiator_needs_a_free_bit(gpu):
    h = device_model(gpu, "real", 65, 64, 6.0, 1.0)   # B = 128: no room for the class bit
    eng = gpu._native.Engine(h.descriptor(), gpu.IsStochasticInteger().descriptor(), 1, 1000, 1000)
    with pytest.raises(gpu.RimuError):
        eng.set_initiator(1, 1.0)
    eng.close()


def test_two_chunk_upload_validation_and_order(gpu):
    M = N = 50
    rng = np.random.default_rng(11)
    ints = wide_addresses(M, N, 500, rng)
    perm = rng.permutation(len(ints))
    keys = O.pack_keys([ints[i] for i in perm], 2)
    vals = np.arange(1, len(ints) + 1, dtype=np.int64)[perm]
    eng = _engine(gpu, "real", M, N, 6.0, "integer", 1)
    eng.upload(keys, vals)                                # unsorted input is sorted on the host
    gk, gv = eng.download()
    assert O.unpack_keys(gk) == ints and np.array_equal(gv, np.arange(1, len(ints) + 1))
    bad_pop = O.pack_keys([ints[0] ^ (1 << 70)], 2)      # one boson more or fewer
    bad_high = O.pack_keys([ints[0] | (1 << 99)], 2)      # a bit at or above B
    for bad in (bad_pop, bad_high):
        k = np.concatenate([bad, O.pack_keys(ints[1:3], 2)]) if O.unpack_keys(bad)[0] < ints[1] \
            else np.concatenate([O.pack_keys(ints[1:3], 2), bad])
        with pytest.raises(gpu.RimuError):
            eng.upload(k, np.ones(3, np.int64))
    eng.upload(keys[:0], vals[:0])
    st = eng.step(1, 0.0, 1e-3)
    assert st.len == 0 and len(eng) == 0
    eng.close()


def test_two_chunk_overlaps_match_oracle(gpu):
    M = N = 50
    rng = np.random.default_rng(21)
    pool = wide_addresses(M, N, 6000, rng)
    li = sorted(rng.choice(len(pool), 4000, replace=False).tolist())
    ri = sorted(rng.choice(len(pool), 3000, replace=False).tolist())
    lk, rk = O.pack_keys([pool[i] for i in li], 2), O.pack_keys([pool[i] for i in ri], 2)
    lv, rv = rng.normal(size=len(li)) * 3.0, rng.integers(-20, 21, size=len(ri)).astype(np.int64)
    rv[rv == 0] = 1
    h = device_model(gpu, "real", M, N, 4.0, 1.0)
    a = gpu._native.Engine(h.descriptor(), gpu.IsDeterministic().descriptor(), 1, 10_000, 1 << 16)
    b = gpu._native.Engine(h.descriptor(), gpu.IsStochasticInteger().descriptor(), 1, 10_000, 1 << 16)
    a.upload(lk, lv)
    b.upload(rk, rv)
    dev = [h, gpu.G2RealCorrelator(0), gpu.G2RealCorrelator(3), gpu.DensityMatrixDiagonal(1),
           gpu.DensityMatrixDiagonal(M)]
    orc = [("ham", oracle_model("real", M, N, 4.0, 1.0)), ("g2", 0), ("g2", 3), ("density", 1),
           ("density", M)]
    got = gpu.observables.overlap_values(gpu.DeviceDVec(a, h, None), gpu.DeviceDVec(b, h, None), dev)
    a.close()
    b.close()
    want = O.overlaps(lk, lv, rk, rv, M, orc)
    scale = np.abs(lv).sum() * np.abs(rv).max() * 10
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-13 * scale)
    assert got[0] != 0 and got[1] != 0


def test_two_chunk_resume_from_state_file_is_bit_exact(gpu, tmp_path):
    """The reference's (50, 50) real-space setup (99-bit addresses) through the package API:
    a run saved with save_state and resumed from load_state continues with the same vectors
    bit for bit."""
    H = gpu.HubbardReal1D(gpu.near_uniform(50, 50), u=6.0)

    def problem(**kw):
        return gpu.ProjectorMonteCarloProblem(
            H, style=gpu.IsDynamicSemistochastic(), time_step=1e-3, random_seed=7,
            shift_strategy=gpu.DoubleLogUpdate(target_walkers=3000), **kw)

    full = gpu.solve(problem(last_step=120))
    fk, fv = full.engine.download()
    fdf = full.df
    full.close()
    part = gpu.solve(problem(last_step=60))
    f = tmp_path / "wide.arrow"
    gpu.save_state(f, part, step=60, shift=part.shift_parameters.shift)
    part.close()
    dv, meta = gpu.load_state(f)
    assert meta["step"] == 60
    resumed = gpu.solve(problem(start_at=dv, shift=meta["shift"], starting_step=60, last_step=120))
    rk, rv = resumed.engine.download()
    rdf = resumed.df
    resumed.close()
    assert fk.shape[1] == 2 and len(rk) > 500
    assert np.array_equal(fk, rk) and np.array_equal(fv.view(np.uint64), rv.view(np.uint64))
    # the resumed run's first pnorm is the loaded vector's norm summed on the host, the
    # uninterrupted run's is the device's sum over merge tiles: the shifts agree to rounding
    np.testing.assert_allclose(fdf["shift"].to_numpy()[60:], rdf["shift"].to_numpy(), rtol=1e-12)
    assert np.array_equal(fdf["len"].to_numpy()[60:], rdf["len"].to_numpy())


@pytest.mark.parametrize("bits", [8, 12, 16, 24, 40, 128])
@pytest.mark.parametrize("case", [("real", 50, 50, "dynamic", 3000, 20.0), ("mom", 40, 30, "dynamic", 1500, 10.0),
                                  ("real", 32, 32, "integer", 3000, 8)],
                         ids=lambda c: f"{c[0]}{c[1]}x{c[2]}-{c[3]}")
def test_top_bit_sort_matches_oracle(gpu, case, bits):
    """Keys above 40 bits are sorted on their top `bits` mixed bits, then runs that agree
    there are ordered (or, past the run limits, the remaining passes run).  Every setting
    gives the oracle's step bit for bit; 8 bits on the momentum case takes the fallback."""
    kind, M, N, style, n, scale = case
    rng = np.random.default_rng(zlib.crc32(repr((case, bits)).encode()))
    if M + N - 1 > 64:
        keys, vals = wide_vector(M, N, n, rng, "int" if style == "integer" else "float", scale)
    else:
        from test_gpu_parity import random_vector
        keys, vals = random_vector(M, N, n, rng, "int" if style == "integer" else "float", scale)
    seed = 0xF1 + bits
    eng = _engine(gpu, kind, M, N, 4.0, style, seed)
    eng.set_sort_bits(bits)
    eng.upload(keys, vals)
    om = oracle_model(kind, M, N, 4.0, 1.0)
    ck, cv = keys, vals
    for step in (1, 2):
        st = eng.step(step, -2.0, 2e-3)
        gk, gv = eng.download()
        ok, ov, ostats = O.step(om, O.style(style), ck, cv, seed=seed, step=step, shift=-2.0,
                                dtau=2e-3, order=1)
        _compare(style, gk, gv, st, ok, ov, ostats)
        ck, cv = gk, gv
    eng.close()
    with pytest.raises(gpu.RimuError):
        eng2 = _engine(gpu, kind, M, N, 4.0, style, seed)
        try:
            eng2.set_sort_bits(4)
        finally:
            eng2.close()


@pytest.mark.parametrize("kind,M,N", [("real", 50, 50), ("real", 32, 32)])
def test_projection_on_mixed_keys(gpu, kind, M, N):
    """ProjectedEnergy terms on keys stored mixed (above 40 bits): vproj = sum pc_i v[pk_i]
    and hproj = sum hc_i v[hk_i] from the device equal the same sums over the downloaded
    vector."""
    rng = np.random.default_rng(M)
    if M + N - 1 > 64:
        keys, vals = wide_vector(M, N, 3000, rng, "float", 5.0)
        ints = O.unpack_keys(keys)
    else:
        from test_gpu_parity import random_vector
        keys, vals = random_vector(M, N, 3000, rng, "float", 5.0)
        ints = [int(x) for x in keys.tolist()]
    kw = 2 if M + N - 1 > 64 else 1
    eng = _engine(gpu, kind, M, N, 4.0, "dynamic", 3)
    eng.upload(keys, vals)
    # projector: some stored addresses and some that are not
    pk = ints[::50] + wide_addresses(M, N, 20, rng)
    pc = rng.normal(size=len(pk))
    hk = ints[7::40]
    hc = rng.normal(size=len(hk))
    eng.set_projector(O.pack_keys(pk, kw), pc, O.pack_keys(hk, kw), hc)
    st = eng.step(1, -2.0, 1e-3)
    gk, gv = eng.download()
    eng.close()
    v = dict(zip(O.unpack_keys(gk), gv.tolist()))
    want_v = sum(c * v.get(k, 0.0) for k, c in zip(pk, pc))
    want_h = sum(c * v.get(k, 0.0) for k, c in zip(hk, hc))
    assert st.vproj == pytest.approx(want_v, rel=1e-12, abs=1e-12)
    assert st.hproj == pytest.approx(want_h, rel=1e-12, abs=1e-12)
    assert want_v != 0 and want_h != 0
