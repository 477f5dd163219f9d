"""GPU parity of HubbardRealSpace (BoseFS, FermiFS and two-component CompositeFS on
CubicGrid geometries) against the oracle, and the reference's energies through the device.

Bars as for the 1D models: matrix elements, addresses and integer values bit-exact,
floating values bit-exact against the oracle's list order; overlaps within 1e-12;
energies of the deterministic projector within 1e-8 of exact diagonalisation.
"""
import math
import zlib

import numpy as np
import pytest

import oracle as O
from test_oracle_realspace import krylov_energy

pytestmark = pytest.mark.gpu

GEO = {"periodic": lambda r, d: r.PeriodicBoundaries(*d), "hardwall": lambda r, d: r.HardwallBoundaries(*d),
       "ladder": lambda r, d: r.LadderBoundaries(*d)}
FOLD = {"periodic": lambda d: (True,) * len(d), "hardwall": lambda d: (False,) * len(d),
        "ladder": lambda d: tuple(i > 0 for i in range(len(d)))}

# name: geometry kind, dims, components, t, u, v
MODELS = {
    "hub2c_4x4": ("periodic", (4, 4), [("fermi", 5), ("fermi", 5)], [1.0, 1.0], [[0, 4], [4, 0]], None),
    "hub2c_3x5_hw": ("hardwall", (3, 5), [("fermi", 4), ("fermi", 3)], [1.0, 0.7], [[0, 2.5], [2.5, 0]], None),
    "fermi_ladder": ("ladder", (2, 6), [("fermi", 5)], [1.3], [[0]], [[0.4, 0.1]]),
    "bose_2d": ("periodic", (4, 4), [("bose", 12)], [1.0], [[6]], None),
    "bose_3d_trap": ("hardwall", (3, 3, 2), [("bose", 8)], [1.0], [[3]], [[0.5, 0.2, 0.1]]),
    "bose_fermi": ("periodic", (3, 4), [("bose", 4), ("fermi", 3)], [1.0, 2.0], [[2, 1.5], [1.5, 0]], None),
    "hub2c_6x6": ("periodic", (6, 6), [("fermi", 12), ("fermi", 12)], [1.0, 1.0], [[0, 4], [4, 0]], None),   # 72 bits
    "hub2c_8x8": ("periodic", (8, 8), [("fermi", 20), ("fermi", 20)], [1.0, 1.0], [[0, 8], [8, 0]], None),   # 128 bits
    "bose_1d_100": ("hardwall", (64,), [("bose", 40)], [1.0], [[4]], None),                                     # 103 bits
    "bose_8x8": ("periodic", (8, 8), [("bose", 64)], [1.0], [[6]], None),                                       # 127 bits
}


def models(rimu, name):
    geo, dims, comps, t, u, v = MODELS[name]
    M = math.prod(dims)
    C = len(comps)
    u = [r[:C] for r in u[:C]]
    om = O.real_space(dims, comps, t=t, u=u, v=v, fold=FOLD[geo](dims))
    addrs = []
    for kind, n in comps:
        addrs.append(rimu.near_uniform(n, M, rimu.FermiFS if kind == "fermi" else None))
    addr = addrs[0] if C == 1 else rimu.CompositeFS(*addrs)
    h = rimu.HubbardRealSpace(addr, geometry=GEO[geo](rimu, dims), t=t, u=u, v=v)
    return h, om, comps, M


def random_keys(om, comps, M, n, rng):
    out = set()
    tries = 0
    while len(out) < n and tries < 50 * n:
        tries += 1
        onrs = []
        for kind, k in comps:
            if kind == "fermi":
                x = [0] * M
                for i in rng.choice(M, k, replace=False):
                    x[i] = 1
            else:
                cuts = np.sort(rng.choice(k + M - 1, M - 1, replace=False))
                x = [int(y) for y in np.diff(np.concatenate([[-1], cuts, [k + M - 1]])) - 1]
            onrs.append(tuple(x))
        out.add(O.rs_key(om, onrs))
    return sorted(out)


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
        assert st.norm == pytest.approx(ostats["norm"], rel=1e-11)
        assert st.spawns == pytest.approx(ostats["spawns"], rel=1e-11)


@pytest.mark.parametrize("name", list(MODELS))
def test_hamiltonian_elements(gpu, name):
    h, om, comps, M = models(gpu, name)
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    ints = random_keys(om, comps, M, 150, rng)
    kw = O.key_words(om)
    assert kw == h.key_words
    addrs = O.pack_keys(ints, kw)
    diag, nod, _, _ = gpu._native.hamiltonian_eval(h.descriptor(), addrs, np.zeros(len(ints), np.int64))
    rows, ch = [], []
    for a, d, n in zip(ints, diag.tolist(), nod.tolist()):
        assert n == O.num_offdiagonals(om, a)
        assert d == O.diagonal_element(om, a)
        rows += [a] * n
        ch += list(range(1, n + 1))
    _, _, na, ov = gpu._native.hamiltonian_eval(h.descriptor(), O.pack_keys(rows, kw), np.array(ch))
    for a, c, b, v in zip(rows, ch, O.unpack_keys(na), ov.tolist()):
        ob, oval = O.get_offdiagonal(om, a, c)
        assert b == ob and v == oval, (a, c)


CASES = [
    # model, style, vector length, value scale, shift, dtau
    ("hub2c_4x4", "integer", 3000, 8, -3.0, 2e-3),
    ("hub2c_4x4", "dynamic", 3000, 10.0, -3.0, 2e-3),
    ("hub2c_3x5_hw", "integer", 2000, 6, -2.0, 5e-3),
    ("hub2c_3x5_hw", "deterministic", 40, 1.0, 0.5, 1e-2),
    ("fermi_ladder", "dynamic", 500, 5.0, 0.0, 1e-2),
    ("bose_2d", "integer", 3000, 8, -5.0, 1e-3),
    ("bose_2d", "dynamic", 3000, 20.0, -5.0, 1e-3),
    ("bose_3d_trap", "dynamic", 2000, 10.0, 2.0, 2e-3),
    ("bose_fermi", "integer", 2000, 6, -1.0, 5e-3),
    ("bose_fermi", "deterministic", 50, 1.0, 0.0, 1e-2),
    ("hub2c_6x6", "dynamic", 3000, 10.0, -10.0, 1e-3),
    ("hub2c_8x8", "integer", 2000, 6, -20.0, 1e-3),
    ("hub2c_8x8", "dynamic", 2000, 10.0, -20.0, 1e-3),
    ("bose_1d_100", "dynamic", 2000, 10.0, -2.0, 2e-3),
    ("bose_8x8", "dynamic", 2000, 10.0, -5.0, 1e-3),     # every spawn channel nonzero: dense records
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}-{c[1]}")
def test_steps_match_oracle(gpu, case):
    name, style, n, scale, shift, dtau = case
    h, om, comps, M = models(gpu, name)
    rng = np.random.default_rng(zlib.crc32(repr(case).encode()))
    ints = random_keys(om, comps, M, n, rng)
    keys = O.pack_keys(ints, O.key_words(om))
    if style == "integer":
        vals = rng.integers(-scale, scale + 1, size=len(ints)).astype(np.int64)
        vals[vals == 0] = 1
    else:
        vals = rng.normal(scale=scale, size=len(ints))
    sty = {"integer": gpu.IsStochasticInteger(), "dynamic": gpu.IsDynamicSemistochastic(),
           "deterministic": gpu.IsDeterministic()}[style]
    seed = 0x5EA + n
    eng = gpu._native.Engine(h.descriptor(), sty.descriptor(), seed, 4_000_000, 1 << 16)
    eng.upload(keys, vals)
    ck, cv = keys, vals
    for step in ((5, 6, 7) if style != "deterministic" else (5, 6)):
        st = eng.step(step, shift, dtau)
        gk, gv = eng.download()
        ok, ov, ostats = O.step(om, O.style(style), ck, cv, seed=seed, step=step, shift=shift,
                                dtau=dtau, order=1)
        _compare(style, gk, gv, st, ok, ov, ostats)
        ck, cv = gk, gv
    eng.close()


@pytest.mark.parametrize("rule", ["initiator", "coherent"])
def test_initiator_steps_match_oracle(gpu, rule):
    h, om, comps, M = models(gpu, "hub2c_4x4")
    rng = np.random.default_rng(7)
    ints = random_keys(om, comps, M, 3000, rng)
    keys = O.pack_keys(ints, 1)
    vals = rng.normal(scale=6.0, size=len(ints))
    eng = gpu._native.Engine(h.descriptor(), gpu.IsDynamicSemistochastic().descriptor(), 3, 4_000_000, 1 << 16)
    eng.set_initiator({"initiator": 1, "coherent": 3}[rule], 2.0)
    eng.upload(keys, vals)
    ck, cv = keys, vals
    for step in (1, 2):
        st = eng.step(step, -3.0, 2e-3)
        gk, gv = eng.download()
        ok, ov, ostats = O.step_initiator(om, O.style("dynamic"), ck, cv, rule=rule, threshold=2.0,
                                          seed=3, step=step, shift=-3.0, dtau=2e-3, order=1)
        _compare("dynamic", gk, gv, st, ok, ov, ostats)
        ck, cv = gk, gv
    eng.close()


@pytest.mark.parametrize("name", ["hub2c_4x4", "bose_fermi", "hub2c_8x8"])
def test_overlaps_match_oracle(gpu, name):
    h, om, comps, M = models(gpu, name)
    rng = np.random.default_rng(11)
    pool = random_keys(om, comps, M, 4000, rng)
    kw = O.key_words(om)
    li = sorted(rng.choice(len(pool), 3000, replace=False).tolist())
    ri = sorted(rng.choice(len(pool), 2500, replace=False).tolist())
    lk, rk = O.pack_keys([pool[i] for i in li], kw), O.pack_keys([pool[i] for i in ri], kw)
    lv, rv = rng.normal(size=len(li)), rng.normal(size=len(ri))
    a = gpu._native.Engine(h.descriptor(), gpu.IsDeterministic().descriptor(), 1, 10_000, 1 << 16)
    b = gpu._native.Engine(h.descriptor(), gpu.IsDeterministic().descriptor(), 1, 10_000, 1 << 16)
    a.upload(lk, lv)
    b.upload(rk, rv)
    got = gpu.observables.overlap_values(gpu.DeviceDVec(a, h, None), gpu.DeviceDVec(b, h, None), [h])
    with pytest.raises(gpu.RimuError):
        gpu.observables.overlap_values(gpu.DeviceDVec(a, h, None), gpu.DeviceDVec(b, h, None),
                                       [gpu.G2RealCorrelator(1)])
    a.close()
    b.close()
    want = O.overlaps(lk, lv, rk, rv, M, [("ham", om)])
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12 * np.abs(lv).sum() * np.abs(rv).max())


def _deterministic_energy(rimu, h, start, steps=1500, dtau=0.02):
    """Power iteration with the deterministic projector through lomc; the Rayleigh quotient
    of the final vector on the device."""
    dv = rimu.DVec((start, 1.0), style=rimu.IsDeterministic())
    # max_length = 2 * target + 100 must hold the whole (dense) vector
    df, sim = rimu.lomc(h, dv, laststep=steps, dtau=dtau,
                        s_strat=rimu.DoubleLogUpdate(target_walkers=100_000))
    assert len(df) == steps
    v = sim.state_vector
    e = rimu.dot(v, h, v) / rimu.dot(v, v)
    sim.close()
    return e


@pytest.mark.parametrize("onr,want", [((1, 1, 1) + (0,) * 13, -8), ((1, 1) + (0,) * 14, -6)])
def test_reference_energies_2d_fermions(gpu, onr, want):
    """test/Hamiltonians.jl:431-437 through the device: 2 and 3 fermions on PeriodicBoundaries(4, 4)."""
    f = gpu.FermiFS(onr)
    h = gpu.HubbardRealSpace(f, geometry=gpu.PeriodicBoundaries(4, 4))
    assert _deterministic_energy(gpu, h, f) == pytest.approx(want, abs=1e-8)


def test_reference_energy_two_component_3x3(gpu):
    """test/Hamiltonians.jl:441-448: CompositeFS(near_uniform(FermiFS{3,9}),
    near_uniform(FermiFS{2,9})), t = [1, 2], u = 0 on PeriodicBoundaries(3, 3): -16; with
    u12 = 1 the device's energy equals the oracle's exact diagonalisation in the start
    vector's sector."""
    a = gpu.CompositeFS(gpu.near_uniform(3, 9, gpu.FermiFS), gpu.near_uniform(2, 9, gpu.FermiFS))
    h0 = gpu.HubbardRealSpace(a, geometry=gpu.PeriodicBoundaries(3, 3), t=[1, 2], u=[[0, 0], [0, 0]])
    assert _deterministic_energy(gpu, h0, a) == pytest.approx(-16, abs=1e-8)
    h1 = gpu.HubbardRealSpace(a, geometry=gpu.PeriodicBoundaries(3, 3), t=[1, 2], u=[[0, 1], [1, 0]])
    om = O.real_space((3, 3), [("fermi", 3), ("fermi", 2)], t=[1, 2], u=[[0, 1], [1, 0]])
    e0 = krylov_energy(om, a.bits)
    # the sector's gap is small: a longer projection
    assert _deterministic_energy(gpu, h1, a, steps=4000, dtau=0.025) == pytest.approx(e0, abs=1e-8)
    assert e0 > -16


def test_upload_rejects_wrong_components(gpu):
    h, om, comps, M = models(gpu, "hub2c_4x4")
    eng = gpu._native.Engine(h.descriptor(), gpu.IsStochasticInteger().descriptor(), 1, 100, 100)
    good = O.rs_key(om, [(1,) * 5 + (0,) * 11, (1,) * 5 + (0,) * 11])
    bad = O.rs_key(om, [(1,) * 6 + (0,) * 10, (1,) * 4 + (0,) * 12])   # 6 + 4 instead of 5 + 5
    eng.upload(np.array([good], np.uint64), np.array([3]))
    with pytest.raises(gpu.RimuError):
        eng.upload(np.array([bad], np.uint64), np.array([3]))
    eng.close()


def test_state_file_round_trip_two_component(gpu, tmp_path):
    """save_state / load_state of a FermiFS2C run (Rimu.CompositeFS keys): the loaded vector
    equals the device vector, and a run resumed from it continues bit for bit."""
    a = gpu.FermiFS2C(16, (1, 1), (2, 1), (1, -1), (2, -1))
    H = gpu.HubbardRealSpace(a, geometry=gpu.PeriodicBoundaries(4, 4), u=[[0, 4], [4, 0]])

    def problem(**kw):
        return gpu.ProjectorMonteCarloProblem(
            H, style=gpu.IsDynamicSemistochastic(), time_step=1e-3, random_seed=5,
            shift_strategy=gpu.DoubleLogUpdate(target_walkers=2000), **kw)

    full = gpu.solve(problem(last_step=80))
    fk, fv = full.engine.download()
    full.close()
    part = gpu.solve(problem(last_step=40))
    pk, pv = part.engine.download()
    f = tmp_path / "fermi2c.arrow"
    gpu.save_state(f, part, step=40, shift=part.shift_parameters.shift)
    part.close()
    dv, meta = gpu.load_state(f)
    assert meta["step"] == 40
    got = sorted((k.bits, v) for k, v in dv.pairs())
    assert [k for k, _ in got] == O.unpack_keys(pk)
    assert np.array_equal(np.array([v for _, v in got]).view(np.uint64), pv.view(np.uint64))
    resumed = gpu.solve(problem(start_at=dv, shift=meta["shift"], starting_step=40, last_step=80))
    rk, rv = resumed.engine.download()
    resumed.close()
    assert len(rk) > 50
    assert np.array_equal(fk, rk) and np.array_equal(fv.view(np.uint64), rv.view(np.uint64))
