"""GPU parity of the step with initiator rules (rimu_engine_set_initiator) against the
oracle's or_step_initiator, and the reference's initiator tests through the package API.

Bars as for the plain step: addresses and integer values bit-exact, floating values
bit-exact against the oracle's list order (each of the three components is a left fold of
its deposits in list order), statistics equal.  Zero values that NoCompression keeps
(addresses reached only by discarded deposits) are part of the compared vectors.
"""
import zlib

import numpy as np
import pytest

import oracle as O
from rimu_amd import statstools as ST
from test_gpu_parity import device_model, oracle_model, random_vector

pytestmark = pytest.mark.gpu

RULE_CODES = {"initiator": 1, "simple": 2, "coherent": 3}

CASES = [
    # model, M, N, u, style, vector length, value scale, shift, dtau, threshold
    ("real", 10, 10, 6.0, "integer", 2000, 6, -2.0, 0.01, 2.0),
    ("mom", 8, 8, 4.0, "integer", 1500, 8, -3.0, 0.005, 3.0),
    ("real", 32, 32, 6.0, "integer", 2000, 6, -2.0, 0.01, 2.0),          # B = 63
    ("real", 20, 20, 6.0, "dynamic", 5000, 60.0, 0.5, 0.001, 20.0),
    ("mom", 8, 8, 4.0, "dynamic", 1500, 10.0, -3.0, 0.005, 5.0),
    ("mom", 20, 10, 1.0, "dynamic", 3000, 5.0, -1.0, 0.01, 1.0),
    ("real", 8, 8, 2.0, "deterministic", 1000, 1.0, 0.3, 0.05, 1.0),
    ("mom", 7, 6, 3.0, "deterministic", 600, 1.0, 0.3, 0.05, 0.8),
]


def _engine(rimu, kind, M, N, u, style, seed, rule, thr):
    h = device_model(rimu, kind, M, N, u, 1.0)
    st = {"integer": rimu.IsStochasticInteger(), "dynamic": rimu.IsDynamicSemistochastic(),
          "deterministic": rimu.IsDeterministic()}[style]
    eng = rimu._native.Engine(h.descriptor(), st.descriptor(), seed, 4_000_000, 1 << 16)
    eng.set_initiator(RULE_CODES[rule], thr)
    return eng


def _compare(style, gk, gv, st, ok, ov, ostats):
    assert np.array_equal(gk, ok)
    if style == "integer":
        assert np.array_equal(gv, ov)
        for name, field in [("attempts", "spawn_attempts"), ("spawns", "spawns"),
                            ("deaths", "deaths"), ("clones", "clones"), ("zombies", "zombies"),
                            ("len", "len"), ("len_before", "len_before"), ("norm", "norm")]:
            assert getattr(st, field) == ostats[name], name
    else:
        # values compared as numbers (a kept zero may carry either sign), then bit for bit
        # where they are nonzero
        assert np.array_equal(gv == 0, ov == 0)
        nz = ov != 0
        assert np.array_equal(gv[nz].view(np.uint64), ov[nz].view(np.uint64))
        assert st.spawn_attempts == ostats["attempts"]
        assert st.exact_steps == ostats["exact_steps"]
        assert st.inexact_steps == ostats["inexact_steps"]
        assert st.len == ostats["len"] and st.len_before == ostats["len_before"]
        assert st.norm == pytest.approx(ostats["norm"], rel=1e-13)
        assert st.spawns == pytest.approx(ostats["spawns"], rel=1e-13)


@pytest.mark.parametrize("rule", ["initiator", "simple", "coherent"])
@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}{c[1]}x{c[2]}-{c[4]}")
def test_initiator_steps_match_oracle(gpu, case, rule):
    kind, M, N, u, style, n, scale, shift, dtau, thr = case
    rng = np.random.default_rng(zlib.crc32(repr((case, rule)).encode()))
    keys, vals = random_vector(M, N, n, rng, "int" if style == "integer" else "float", scale)
    seed = 0xB0B + n
    eng = _engine(gpu, kind, M, N, u, style, seed, rule, thr)
    eng.upload(keys, vals)
    om = oracle_model(kind, M, N, u, 1.0)
    osty = O.style(style)
    ck, cv = keys, vals
    zeros_seen = 0
    for step in (11, 12, 13):
        st = eng.step(step, shift, dtau)
        gk, gv = eng.download()
        ok, ov, ostats = O.step_initiator(om, osty, ck, cv, rule=rule, threshold=thr, seed=seed,
                                          step=step, shift=shift, dtau=dtau, order=1)
        _compare(style, gk, gv, st, ok, ov, ostats)
        zeros_seen += int((gv == 0).sum())
        ck, cv = gk, gv   # the next step starts from the engine's own vector (zeros included)
    eng.close()
    if style != "dynamic":
        assert zeros_seen > 0   # NoCompression keeps discarded addresses as zero entries


def test_initiator_runs_crossing_merge_tiles(gpu):
    """A few entries with thousands of walkers: runs of one address span several merge
    tiles, so the fold reads staged and global records past its tile."""
    M = N = 6
    rng = np.random.default_rng(3)
    keys = np.sort(rng.choice(O.build_basis(M, N), 24, replace=False))
    vals = rng.integers(1, 6000, size=24).astype(np.int64) * rng.choice([-1, 1], size=24)
    vals[::3] = rng.integers(-2, 3, size=8)    # non-initiators among them
    vals[vals == 0] = 1
    om = oracle_model("real", M, N, 6.0, 1.0)
    for rule in ("initiator", "coherent"):
        eng = _engine(gpu, "real", M, N, 6.0, "integer", 77, rule, 2.0)
        eng.upload(keys, vals)
        st = eng.step(5, 1.0, 0.05)
        gk, gv = eng.download()
        eng.close()
        ok, ov, ostats = O.step_initiator(om, O.style("integer"), keys, vals, rule=rule,
                                          threshold=2.0, seed=77, step=5, shift=1.0, dtau=0.05,
                                          order=1)
        assert st.spawn_attempts > 20000
        _compare("integer", gk, gv, st, ok, ov, ostats)


def test_initiator_rule_off_is_the_plain_step(gpu):
    M = N = 10
    rng = np.random.default_rng(4)
    keys, vals = random_vector(M, N, 3000, rng, "float", 8.0)
    outs = []
    for rule in (None, 0):
        h = device_model(gpu, "real", M, N, 6.0, 1.0)
        eng = gpu._native.Engine(h.descriptor(), gpu.IsDynamicSemistochastic().descriptor(), 9,
                                 1_000_000, 1 << 16)
        if rule is not None:
            eng.set_initiator(2, 1.0)
            eng.set_initiator(rule, 1.0)
        eng.upload(keys, vals)
        eng.step(1, -1.0, 0.01)
        outs.append(eng.download())
        eng.close()
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1].view(np.uint64), outs[1][1].view(np.uint64))


def test_initiator_argument_errors(gpu):
    h = device_model(gpu, "real", 33, 32, 6.0, 1.0)   # B = 64: no room for the class bit
    eng = gpu._native.Engine(h.descriptor(), gpu.IsStochasticInteger().descriptor(), 1, 1000, 1000)
    with pytest.raises(gpu.RimuError):
        eng.set_initiator(1, 1.0)
    eng.set_initiator(0, 1.0)
    with pytest.raises(gpu.RimuError):
        eng.set_initiator(4, 1.0)
    eng.close()
    h = device_model(gpu, "real", 8, 8, 6.0, 1.0)
    eng = gpu._native.Engine(h.descriptor(), gpu.IsStochasticInteger().descriptor(), 1, 1000, 1000)
    with pytest.raises(gpu.RimuError):
        eng.set_initiator(1, float("nan"))
    eng.close()


def test_initiator_energies_below_the_plateau(gpu):
    """test/lomc.jl:612-675 through lomc with InitiatorDVec start vectors: BoseFS{10,10}
    condensate, HubbardMom1D u = 4, 300 walkers, dtau = 5e-4, 6000 steps."""
    address = gpu.BoseFS(0, 0, 0, 0, 10, 0, 0, 0, 0, 0)
    H = gpu.HubbardMom1D(address, u=4.0)
    E0 = -9.251592973178997
    s_strat = gpu.DoubleLogUpdate(zeta=0.05, xi=0.05 ** 2 / 4, target_walkers=300)
    sty = gpu.IsDynamicSemistochastic()
    vecs = {
        "no": gpu.DVec((address, 1), style=sty),
        "i1": gpu.InitiatorDVec((address, 1), initiator=gpu.Initiator(1), style=sty),
        "i2": gpu.InitiatorDVec((address, 1), initiator=gpu.SimpleInitiator(1), style=sty),
        "i3": gpu.InitiatorDVec((address, 1), initiator=gpu.CoherentInitiator(1), style=sty),
        "ni": gpu.InitiatorDVec((address, 1), initiator=gpu.NonInitiator(), style=sty),
    }
    est = {}
    for name, dv in vecs.items():
        df, sim = gpu.lomc(H, dv.copy(), s_strat=s_strat, laststep=6000, dtau=5e-4,
                           random_seed=8008)
        assert sim.v.initiator == dv.initiator
        sim.close()
        br = ST.blocking_analysis(df["shift"].to_numpy()[1999:])
        est[name] = (br.mean, br.err)
    (E_no, s_no), (E_ni, _) = est["no"], est["ni"]
    (E_i1, s_i1), (E_i2, _), (E_i3, s_i3) = est["i1"], est["i2"], est["i3"]
    assert E_no < E0 and E_ni < E0
    assert abs(E_no - E_ni) <= 3 * s_no
    assert E_i1 > E0 and E_i2 > E0 and E_i3 > E0
    assert E_i2 > E_i1
    assert abs(E_i1 - E_i3) <= max(3 * s_i1, 3 * s_i3)


def test_integer_initiator_fciqmc_shift_above_e0(gpu):
    """test/mpi_runtests.jl:190-203: HubbardMom1D(BoseFS(0,0,0,7,0,0,0); u = 6), start
    PDVec(addr => 3; initiator_threshold = 1) (integer walkers, Initiator(1)),
    DoubleLogUpdate(target_walkers = 100), 5000 steps: the initiator shift lies at or above
    E0 = -4.628524493494574."""
    addr = gpu.BoseFS(0, 0, 0, 7, 0, 0, 0)
    H = gpu.HubbardMom1D(addr, u=6.0)
    E0 = -4.628524493494574
    dv = gpu.PDVec((addr, 3), initiator_threshold=1)
    assert dv.initiator == gpu.Initiator(1) and isinstance(dv.style, gpu.IsStochasticInteger)
    p = gpu.ProjectorMonteCarloProblem(H, start_at=dv, last_step=5000, random_seed=13,
                                       shift_strategy=gpu.DoubleLogUpdate(target_walkers=100))
    sim = gpu.solve(p)
    df = sim.df
    sim.close()
    assert len(df) == 5000
    Es = ST.blocking_analysis(df["shift"].to_numpy()[1999:]).mean
    assert E0 <= Es
