"""DoubleLogUpdate on the device (rimu_engine_set_shift_update) through the C ABI: every
step's shift against the closed form of update_shift_parameters!
(shiftstrategy.jl:174-181) applied to the step's own norm, the step reading the device
value (bit-identical to a run that passes the same shifts from the host), repeated and
failed steps leaving the shift parameters unchanged, and the pmc loop on the device update
against the host strategy."""
import numpy as np
import pytest

import oracle as O
from test_gpu_parity import device_model, random_vector

pytestmark = pytest.mark.gpu


def _engine(rimu, kind, M, N, u, style, seed, max_spawns=1 << 20, mode=2):
    h = device_model(rimu, kind, M, N, u, 1.0)
    st = {"integer": rimu.IsStochasticInteger(), "dynamic": rimu.IsDynamicSemistochastic()}[style]
    e = rimu._native.Engine(h.descriptor(), st.descriptor(), seed, 4_000_000, max_spawns)
    e.set_radix_mode(mode)
    return e


def _check_trajectory(shifts, norms, s0, n0, target, zeta, xi, dtau):
    prev_s, prev_n = s0, n0
    for i, (s, n) in enumerate(zip(shifts, norms)):
        want = O.double_log_update(prev_s, n, prev_n, target, zeta=zeta, xi=xi, dtau=dtau)
        assert s == pytest.approx(want, rel=1e-12, abs=1e-12), i
        prev_s, prev_n = s, n


CASES = [  # kind, M, N, u, style, start population, target, zeta, xi, dtau, steps
    ("real", 10, 10, 6.0, "integer", 10, 2000, 0.08, None, 1e-2, 400),
    ("real", 12, 12, 4.0, "dynamic", 20.0, 5000, 0.3, 0.01, 5e-3, 300),
    ("mom", 9, 9, 3.0, "dynamic", 30.0, 3000, 0.05, 0.05 ** 2 / 4, 1e-2, 300),
]


@pytest.mark.parametrize("kind,M,N,u,style,pop,target,zeta,xi,dtau,steps", CASES)
def test_device_shift_follows_double_log_update(gpu, kind, M, N, u, style, pop, target, zeta, xi,
                                                dtau, steps):
    """Every step's shift equals DoubleLogUpdate of the step's norm (to 1e-12: the device's
    and the host's ln may differ in the last place); pnorm is the norm; the population
    settles at the target."""
    print(f"device shift {kind} {M}x{N} {style} target {target}", flush=True)
    h = device_model(gpu, kind, M, N, u, 1.0)
    e = _engine(gpu, kind, M, N, u, style, 41 + M)
    start = h.starting_address().bits
    e.upload(np.array([start], np.uint64), np.array([pop]))
    s0 = h.diagonal_element(h.starting_address())
    xi_v = zeta ** 2 / 4 if xi is None else xi
    e.set_shift_update(gpu._native.SHIFT_DOUBLE_LOG, s0, abs(pop), target, zeta, xi)
    shifts, norms = [], []
    for step in range(1, steps + 1):
        st = e.step(step, float("nan"), dtau)   # the argument is ignored
        s, pn = e.shift_parameters()
        assert pn == st.norm
        shifts.append(s)
        norms.append(st.norm)
    e.close()
    assert np.isfinite(shifts).all()
    _check_trajectory(shifts, norms, s0, float(abs(pop)), target, zeta, xi_v, dtau)
    assert 0.25 * target < np.mean(norms[-50:]) < 4 * target


def test_device_shift_is_the_step_shift(gpu):
    """A run on the device update and a run given the same shifts from the host give the
    same vectors and statistics step by step: the step reads the device value."""
    M = N = 12
    h = device_model(gpu, "real", M, N, 4.0, 1.0)
    start = np.array([h.starting_address().bits], np.uint64)
    s0 = h.diagonal_element(h.starting_address())
    a = _engine(gpu, "real", M, N, 4.0, "dynamic", 5)
    b = _engine(gpu, "real", M, N, 4.0, "dynamic", 5)
    a.upload(start, np.array([10.0]))
    b.upload(start, np.array([10.0]))
    a.set_shift_update(gpu._native.SHIFT_DOUBLE_LOG, s0, 10.0, 3000, 0.08, None)
    shift = s0
    for step in range(1, 201):
        sa = a.step(step, 0.0, 1e-2)
        sb = b.step(step, shift, 1e-2)
        assert (sa.len, sa.norm, sa.spawns, sa.deaths) == (sb.len, sb.norm, sb.spawns, sb.deaths), step
        shift = a.shift_parameters()[0]
    ka, va = a.download()
    kb, vb = b.download()
    a.close()
    b.close()
    assert len(ka) > 500
    assert np.array_equal(ka, kb) and np.array_equal(va.view(np.uint64), vb.view(np.uint64))


def test_repeated_and_failed_steps_keep_the_shift(gpu):
    """A step repeated after a forced look-back give-up, and a step that overflowed the spawn
    buffer (the wrapper grows it and repeats the step), update the shift once: the
    trajectory still follows DoubleLogUpdate of the kept steps' norms."""
    M = N = 20
    rng = np.random.default_rng(12)
    keys, vals = random_vector(M, N, 40_000, rng, "float", 30.0)
    n0 = float(np.abs(vals).sum())
    target, zeta, dtau = 2 * n0, 0.1, 1e-3
    e = _engine(gpu, "real", M, N, 6.0, "dynamic", 9, max_spawns=4096)
    e.upload(keys, vals)
    e.set_shift_update(gpu._native.SHIFT_DOUBLE_LOG, -1.0, n0, target, zeta, None)
    e.set_timing(True)
    shifts, norms = [], []
    for step in range(1, 7):
        if step == 3:
            e.set_lookback(1000, withhold_first=True)
        st = e.step(step, 0.0, dtau)
        if step == 3:
            e.set_lookback(1 << 25, withhold_first=False)
            assert "lookback_fallback" in e.kernel_times()
        shifts.append(e.shift_parameters()[0])
        norms.append(st.norm)
    assert e.max_spawns > 4096   # the first step overflowed and was repeated
    e.close()
    _check_trajectory(shifts, norms, -1.0, n0, target, zeta, zeta ** 2 / 4, dtau)


def test_pmc_device_shift_matches_host_strategy(gpu):
    """solve() with DoubleLogUpdate (device update) against the same problem with the same
    update on the host: the same report columns, the same shifts while the trajectories
    agree, and the device run's shifts follow DoubleLogUpdate of its norms."""
    H = gpu.HubbardReal1D(gpu.near_uniform(10, 10), u=6.0)

    class HostDoubleLog(gpu.DoubleLogUpdate):
        """The same update on the host (a subclass is not run on the device)."""

    out = []
    for strat in (gpu.DoubleLogUpdate(target_walkers=1000), HostDoubleLog(target_walkers=1000)):
        p = gpu.ProjectorMonteCarloProblem(H, last_step=300, shift_strategy=strat, random_seed=3,
                                           style=gpu.IsStochasticInteger())
        sim = gpu.solve(p)
        assert sim.device_shift == (type(strat) is gpu.DoubleLogUpdate)
        out.append(sim.df)
        sim.close()
    dev, host = out
    assert list(dev.columns) == list(host.columns)
    # the first steps agree exactly or to the last place of ln; later the populations may
    # part when a stochastic rounding sees another last bit of the shift
    np.testing.assert_allclose(dev["shift"].to_numpy()[:20], host["shift"].to_numpy()[:20], rtol=1e-12)
    # the default initial shift is the Rayleigh quotient of the one-address start vector
    s0 = H.diagonal_element(H.starting_address())
    _check_trajectory(dev["shift"].to_numpy(), dev["norm"].to_numpy(), s0, 10.0, 1000, 0.08,
                      0.08 ** 2 / 4, 0.01)


@pytest.mark.parametrize("case", ["wide", "initiator"])
def test_pmc_device_shift_two_chunks_and_initiator(gpu, case):
    """solve() on two-chunk addresses (HubbardReal1D 50x50, 99 bits) and with Initiator(1)
    (HubbardMom1D): the report's shifts follow DoubleLogUpdate of its norms step by step."""
    if case == "wide":
        H = gpu.HubbardReal1D(gpu.near_uniform(50, 50), u=6.0)
        kw = dict(style=gpu.IsDynamicSemistochastic(), initiator=False)
    else:
        H = gpu.HubbardMom1D(gpu.near_uniform(10, 12), u=2.0)
        kw = dict(style=gpu.IsDynamicSemistochastic(), initiator=True)
    strat = gpu.DoubleLogUpdate(target_walkers=2000, zeta=0.1)
    p = gpu.ProjectorMonteCarloProblem(H, time_step=1e-3, last_step=400, random_seed=11,
                                       shift_strategy=strat, **kw)
    sim = gpu.solve(p)
    assert sim.device_shift and sim.success
    df = sim.df
    sim.close()
    s0 = H.diagonal_element(H.starting_address())   # one-address start: its Rayleigh quotient
    _check_trajectory(df["shift"].to_numpy(), df["norm"].to_numpy(), s0, 10.0, 2000, 0.1,
                      0.1 ** 2 / 4, 1e-3)
    assert df["norm"].to_numpy()[-1] > 200
