"""ProjectorMonteCarloProblem / solve / lomc on the plain engine: the reference's abort
cases (test/projector_monte_carlo_problem.jl:211-312), the report, seeding, and energies
against exact diagonalisation."""
import math

import numpy as np
import pytest

import oracle as O
from rimu_amd import statstools as ST

pytestmark = pytest.mark.gpu


def test_dead_population_aborts(gpu):
    address = gpu.BoseFS(2, 3)
    H = gpu.HubbardReal1D(address, u=20)
    dv = gpu.DVec((address, 10), style=gpu.IsStochasticInteger())
    p = gpu.ProjectorMonteCarloProblem(H, start_at=dv, last_step=100, shift=0.0, random_seed=7)
    sim = gpu.solve(p)
    assert sim.aborted and not sim.success and sim.modified
    assert sim.message.startswith("Aborted in step ") and sim.step < 100
    assert len(sim.df) < 100
    # the population does not die with the default shift (Rayleigh quotient of the start)
    p = gpu.ProjectorMonteCarloProblem(H, start_at=dv, last_step=100, random_seed=7)
    sim = gpu.solve(p)
    assert sim.success and not sim.aborted and len(sim.df) == 100


def test_max_length_aborts(gpu):
    h = gpu.HubbardReal1D(gpu.BoseFS(1, 3, 5, 2, 1))
    p = gpu.ProjectorMonteCarloProblem(h, time_step=0.1, target_walkers=100, random_seed=7)
    sim = gpu.init(p)
    assert sim.problem.max_length == 300 and sim.step == 0
    sim.solve_()
    assert sim.aborted and not sim.success and sim.message.startswith("Aborted in step ")
    step_before = sim.step
    sim.step_()  # no effect after an abort
    assert sim.step == step_before
    p = gpu.ProjectorMonteCarloProblem(h, time_step=0.01, target_walkers=100, random_seed=7)
    sim = gpu.solve(p)
    assert sim.success and len(sim.df) == 100


def test_wall_time_aborts(gpu):
    h = gpu.HubbardReal1D(gpu.BoseFS(1, 3, 5, 2, 1))
    p = gpu.ProjectorMonteCarloProblem(h, last_step=10_000, wall_time=1e-3, random_seed=3)
    sim = gpu.solve(p)
    assert sim.aborted and sim.message == "Wall time limit reached."


def test_report_columns_and_seeding(gpu):
    h = gpu.HubbardReal1D(gpu.BoseFS(1, 2, 3))
    p = gpu.ProjectorMonteCarloProblem(h, target_walkers=500, last_step=100, random_seed=17,
                                       post_step_strategy=gpu.ProjectedEnergy(h, h.starting_address()))
    a = gpu.solve(p).df
    b = gpu.solve(p).df
    assert list(a.columns) == ["step", "len", "shift", "norm", "exact_steps", "inexact_steps",
                               "spawn_attempts", "spawns", "len_before", "vproj", "hproj"]
    assert len(a) == 100
    np.testing.assert_array_equal(a.to_numpy(), b.to_numpy())   # same seed, same run
    df, sim = gpu.lomc(h, laststep=50)
    assert list(df.columns)[:4] == ["step", "len", "shift", "norm"]
    assert "zombies" in df.columns and len(df) == 50


def test_stochastic_style_comparison(gpu):
    """test/lomc.jl:543-610 for the styles this engine has: HubbardReal1D 5x5,
    E0 = -8.280991746582686, 2500 steps; report columns per style, the noise ordering
    sigma(integer) > sigma(dynamic), and every shift within 3 sigma of E0."""
    address = gpu.BoseFS(1, 1, 1, 1, 1)
    H = gpu.HubbardReal1D(address)
    E0 = -8.280991746582686
    styles = {
        "st": gpu.IsStochasticInteger(),
        "dy": gpu.IsDynamicSemistochastic(),
        "de": gpu.IsDeterministic(),
        "dp": gpu.IsDeterministic(gpu.ThresholdCompression()),
    }
    dfs = {}
    for i, (name, st) in enumerate(styles.items()):
        s_strat = gpu.DoubleLogUpdate(zeta=0.05, xi=0.05 ** 2 / 4, target_walkers=100)
        dfs[name], _ = gpu.lomc(H, gpu.DVec((address, 1), style=st), s_strat=s_strat,
                                laststep=2500, random_seed=1234 + i)
    cols = {k: set(df.columns) for k, df in dfs.items()}
    assert {"spawns", "deaths", "clones", "zombies"} <= cols["st"]
    assert {"exact_steps", "inexact_steps", "spawns"} <= cols["dy"]
    assert "exact_steps" in cols["de"]
    assert {"exact_steps", "len_before"} <= cols["dp"]
    assert "len_before" not in cols["st"] and "len_before" not in cols["de"]
    dp = dfs["dp"]
    assert (dp["len_before"] > 0).all() and (dp["len_before"] >= dp["len"]).all()
    res = {k: ST.blocking_analysis(df["shift"].to_numpy()[499:]) for k, df in dfs.items()}
    assert res["st"].err > res["dy"].err
    for k, r in res.items():
        assert abs(r.mean - E0) < 3 * r.err, (k, r)


def test_momentum_energy_above_the_plateau(gpu):
    """test/lomc.jl:675-702 without initiators: HubbardMom1D 10x10 condensate, u = 1,
    E0 = -16.36048582876015, 3000 walkers, dtau = 0.01, 2500 steps."""
    address = gpu.BoseFS(0, 0, 0, 0, 10, 0, 0, 0, 0, 0)
    H = gpu.HubbardMom1D(address)
    E0 = -16.36048582876015
    s_strat = gpu.DoubleLogUpdate(zeta=0.05, xi=0.05 ** 2 / 4, target_walkers=3000)
    dv = gpu.DVec((address, 1), style=gpu.IsDynamicSemistochastic())
    df, _ = gpu.lomc(H, dv, s_strat=s_strat, laststep=2500, dtau=1e-2, random_seed=1337)
    r = ST.blocking_analysis(df["shift"].to_numpy()[499:])
    assert abs(r.mean - E0) < 3 * r.err, r


def test_momentum_energy_below_the_plateau(gpu):
    """test/lomc.jl:633-663 without initiators: u = 4, E0 = -9.251592973178997, 300 walkers
    sit below the plateau, so the shift comes out below E0 (the sign problem's bias)."""
    address = gpu.BoseFS(0, 0, 0, 0, 10, 0, 0, 0, 0, 0)
    H = gpu.HubbardMom1D(address, u=4.0)
    E0 = -9.251592973178997
    s_strat = gpu.DoubleLogUpdate(zeta=0.05, xi=0.05 ** 2 / 4, target_walkers=300)
    dv = gpu.DVec((address, 1), style=gpu.IsDynamicSemistochastic())
    df, _ = gpu.lomc(H, dv, s_strat=s_strat, laststep=6000, dtau=5e-4, random_seed=8008)
    r = ST.blocking_analysis(df["shift"].to_numpy()[1999:])
    assert r.mean < E0, r


def test_projected_energy_known_answer(gpu):
    """ProjectedEnergy through ProjectorMonteCarloProblem on HubbardReal1D 5x5: the ratio of
    means of hproj and vproj within 3 sigma of E0 = -8.280991746582686 (test/lomc.jl:546)."""
    address = gpu.BoseFS(1, 1, 1, 1, 1)
    H = gpu.HubbardReal1D(address)
    E0 = -8.280991746582686
    p = gpu.ProjectorMonteCarloProblem(
        H, start_at=address, time_step=1e-2, last_step=6000, random_seed=4321,
        shift_strategy=gpu.DoubleLogUpdate(target_walkers=300, zeta=0.05, xi=0.05 ** 2 / 4),
        post_step_strategy=gpu.ProjectedEnergy(H, address))
    df = gpu.solve(p).df
    pe = ST.projected_energy(df, skip=1000)
    se = ST.shift_estimator(df, skip=1000)
    assert ST.within(pe, E0, n=3), pe
    assert abs(se.mean - E0) < 3 * se.err, se


def test_config1_energy_against_ed(gpu):
    """HubbardReal1D M = N = 10, u = 6, IsStochasticInteger, 1000 walkers, dtau = 1e-3,
    20,000 steps, seed 17, lomc! defaults (start population 10, shift H_aa = 0): the
    projected energy within 2 sigma of exact diagonalisation (dimension 92,378).  The shift
    of a 1000-walker run carries population-control bias of order 0.1 here (it shrinks with
    the walker number), so the shift is held to 2 sigma at 10,000 walkers."""
    address = gpu.near_uniform(10, 10)
    H = gpu.HubbardReal1D(address, u=6.0)
    E0 = O.ed_energies(O.real1d(10, 10, 6.0, 1.0), O.build_basis(10, 10), 1)[0]
    d = gpu.DenseHamiltonians([H])
    e_lanczos, _ = d.ground_state()
    d.close()
    assert e_lanczos[0] == pytest.approx(E0, abs=1e-9)
    df, _ = gpu.lomc(H, targetwalkers=1000, laststep=20_000, dtau=1e-3, random_seed=17,
                     post_step=gpu.ProjectedEnergy(H, address))
    pe = ST.projected_energy(df, skip=5000)
    assert ST.within(pe, E0, n=2), pe
    df, _ = gpu.lomc(H, targetwalkers=10_000, laststep=20_000, dtau=1e-3, random_seed=17,
                     post_step=gpu.ProjectedEnergy(H, address))
    se = ST.shift_estimator(df, skip=5000)
    pe = ST.projected_energy(df, skip=5000)
    assert abs(se.mean - E0) < 2 * se.err, se
    assert ST.within(pe, E0, n=2), pe
