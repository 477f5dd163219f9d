"""Batched engine: parity with the plain engine, invariance under launch configuration and
graph capture, autotuning that leaves the state alone, and energies against ED."""
import numpy as np
import pytest

import oracle as O
from rimu_amd import statstools as ST

pytestmark = pytest.mark.gpu


def specs(rimu, cases, zeta=0.08, target=500, dtau=1e-3, max_length=None):
    out = []
    for i, (M, N, u) in enumerate(cases):
        h = rimu.HubbardReal1D(rimu.near_uniform(N, M), u=u, t=1.0)
        xi = 0.0 if zeta == 0 else None
        out.append(rimu.RunSpec(h, target_walkers=target, dtau=dtau, seed=100 + i, zeta=zeta, xi=xi,
                                max_length=max_length))
    return out


CASES = [(6, 6, 4.0), (9, 5, 2.5), (12, 12, 8.0), (7, 4, 1.0)]


def test_batch_matches_plain_engine_frozen_shift(gpu):
    """With zeta = xi = 0 the shift stays fixed, so every step of a batched run must equal
    the plain engine's step from the same state (integers: bit-exact)."""
    sp = specs(gpu, CASES, zeta=0.0, dtau=0.01, max_length=200_000)
    nsteps = 60
    b = gpu.BatchedRuns(sp, history_steps=nsteps)
    b.step(nsteps, steps_per_graph=1)
    hist = b.history()
    for r, s in enumerate(sp):
        h = s.hamiltonian
        d = s.descriptor()
        eng = gpu._native.Engine(h.descriptor(), gpu.IsStochasticInteger().descriptor(), d.seed,
                                 d.max_length + 1, 1 << 16)
        eng.upload(np.array([d.start_address], np.uint64), np.array([d.start_population]))
        norms = []
        for step in range(1, nsteps + 1):
            st = eng.step(step, d.shift, d.dtau)
            norms.append(st.norm)
        k, v = eng.download()
        kb, vb = b.vector(r)
        assert np.array_equal(k, kb) and np.array_equal(v, vb), f"run {r}"
        np.testing.assert_array_equal(hist["norm"][:, r], norms)
        eng.close()
    b.close()


def test_batch_config_and_graph_invariance(gpu):
    sp = specs(gpu, CASES, target=2000)
    results = []
    for blk, rad, spg in [(512, 8, 25), (256, 4, 1), (1024, 6, 7)]:
        b = gpu.BatchedRuns(sp, history_steps=200)
        b.set_config(blk, rad)
        b.step(200, steps_per_graph=spg)
        h = b.history()
        vecs = [b.vector(r) for r in range(len(sp))]
        results.append((h, vecs))
        b.close()
    h0, v0 = results[0]
    for h, v in results[1:]:
        for f in ("shift", "norm", "len", "vproj", "hproj"):
            np.testing.assert_array_equal(h[f], h0[f])
        for (ka, va), (kb, vb) in zip(v, v0):
            assert np.array_equal(ka, kb) and np.array_equal(va, vb)


def test_autotune_keeps_state(gpu):
    sp = specs(gpu, CASES, target=2000)
    a = gpu.BatchedRuns(sp, history_steps=300)
    a.step(150, steps_per_graph=50)
    blk, rad, table = a.autotune(trial_steps=5)
    assert blk in (256, 512, 1024) and rad in (4, 6, 8)
    assert len(table) == 9 and all(t > 0 for t in table.values())
    a.step(150, steps_per_graph=50)
    b = gpu.BatchedRuns(sp, history_steps=300)
    b.step(300, steps_per_graph=50)
    ha, hb = a.history(), b.history()
    np.testing.assert_array_equal(ha["norm"], hb["norm"])
    np.testing.assert_array_equal(ha["shift"], hb["shift"])


def test_batch_status_dead_population_and_max_length(gpu):
    """The reference's abort set-ups (test/projector_monte_carlo_problem.jl:252-312): a dead
    population with BoseFS(2,3), u = 20, shift 0; a max_length overflow with
    BoseFS(1,3,5,2,1), dtau = 0.1, 100 target walkers.  Each run stops on its own; the
    healthy run beside them continues."""
    dead = gpu.RunSpec(gpu.HubbardReal1D(gpu.BoseFS(2, 3), u=20.0), target_walkers=1000,
                       dtau=0.01, seed=7, shift=0.0, zeta=0.0, xi=0.0)
    blow = gpu.RunSpec(gpu.HubbardReal1D(gpu.BoseFS(1, 3, 5, 2, 1)), target_walkers=100,
                       dtau=0.1, seed=7)
    ok = gpu.RunSpec(gpu.HubbardReal1D(gpu.BoseFS(2, 3), u=20.0), target_walkers=1000,
                     dtau=0.01, seed=7)
    b = gpu.BatchedRuns([dead, blow, ok], history_steps=100)
    b.step(100, steps_per_graph=10)
    st, ln = b.status()
    assert st[0] == 1 and st[1] == 2 and st[2] == 0 and ln[2] > 0
    h = b.history()
    assert h["norm"].shape[0] == 100


@pytest.mark.parametrize("M,N,u", [(6, 6, 6.0), (8, 5, 3.0), (7, 7, 2.0)])
def test_batch_energy_against_ed(gpu, M, N, u):
    """Shift and projected energy within 3 sigma (blocking errors) of exact diagonalisation."""
    h = gpu.HubbardReal1D(gpu.near_uniform(N, M), u=u)
    nsteps, skip = 30000, 5000
    sp = [gpu.RunSpec(h, target_walkers=3000, dtau=1e-3, seed=7 + s) for s in range(4)]
    b = gpu.BatchedRuns(sp, history_steps=nsteps)
    b.step(nsteps, steps_per_graph=500)
    hist = b.history()
    assert (b.status()[0] == 0).all()
    E0 = O.ed_energies(O.real1d(M, N, u, 1.0), O.build_basis(M, N), 1)[0]
    for r in range(len(sp)):
        se = ST.blocking_analysis(hist["shift"][:, r], skip=skip)
        pe = ST.ratio_of_means(hist["hproj"][:, r], hist["vproj"][:, r], skip=skip)
        assert se.k > 0 and pe.k > 0
        assert abs(se.mean - E0) < 3 * se.err + 1e-3 * abs(E0), (r, se, E0)
        assert abs(pe.ratio - E0) < 3 * pe.err + 1e-3 * abs(E0), (r, pe, E0)
