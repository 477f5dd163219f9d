"""Overlaps on the device (rimu_overlaps) against the oracle's or_overlaps, and replica runs
with AllOverlaps through ProjectorMonteCarloProblem.

Bars: integer dots exact; everything else within 1e-12 relative (the sums run in another
order than the oracle's), operators being Hamiltonians (off-diagonals included),
G2RealCorrelator and DensityMatrixDiagonal.
"""
import math
import zlib

import numpy as np
import pytest

import oracle as O
from rimu_amd import statstools as ST
from test_gpu_parity import device_model, oracle_model, random_addresses

pytestmark = pytest.mark.gpu


def _engine(rimu, kind, M, N, vtype, keys, vals):
    h = device_model(rimu, kind, M, N, 4.0, 1.0)
    st = rimu.IsStochasticInteger() if vtype == "int" else rimu.IsDeterministic()
    eng = rimu._native.Engine(h.descriptor(), st.descriptor(), 1, max(len(keys), 1) + 10, 1 << 16)
    eng.upload(keys, vals)
    return eng


def _vector(M, N, n, rng, vtype, pool=None):
    keys = random_addresses(M, N, n, rng) if pool is None else np.sort(rng.choice(pool, n, replace=False))
    vals = (rng.integers(-20, 21, size=len(keys)).astype(np.int64) if vtype == "int"
            else rng.normal(size=len(keys)) * 3.0)
    vals[vals == 0] = 1
    return keys, vals


def _ops(rimu, kind, M, N):
    h1 = device_model(rimu, kind, M, N, 4.0, 1.0)
    h2 = device_model(rimu, kind, M, N, 1.5, 0.7)
    dev = [h1, h2, rimu.G2RealCorrelator(0), rimu.G2RealCorrelator(1),
           rimu.G2RealCorrelator(-2), rimu.DensityMatrixDiagonal(1), rimu.DensityMatrixDiagonal(M)]
    orc = [("ham", oracle_model(kind, M, N, 4.0, 1.0)), ("ham", oracle_model(kind, M, N, 1.5, 0.7)),
           ("g2", 0), ("g2", 1), ("g2", -2), ("density", 1), ("density", M)]
    return dev, orc


CASES = [
    # model, M, N, left length, right length, left values, right values, shared pool
    ("real", 8, 8, 3000, 2000, "float", "float", True),
    ("real", 8, 8, 500, 4000, "int", "int", True),
    ("real", 20, 20, 20000, 30000, "float", "int", False),
    ("mom", 8, 6, 1500, 1200, "int", "float", True),
    ("mom", 12, 12, 4000, 4000, "float", "float", False),
    ("real", 32, 32, 3000, 3000, "int", "int", False),
    # mixed keys (23 bits) and a left vector of more than 2^16 entries: neighbours are looked
    # up through the top-bit index of the left keys (6% of the space: ~1.4 hits per entry)
    ("real", 12, 12, 80000, 70000, "float", "float", False),
    ("mom", 12, 12, 90000, 50000, "int", "float", False),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}{c[1]}x{c[2]}-{c[3]}-{c[4]}")
def test_overlaps_match_oracle(gpu, case):
    kind, M, N, nl, nr, vl, vr, shared = case
    rng = np.random.default_rng(zlib.crc32(repr(case).encode()))
    pool = None
    if shared:
        pool = random_addresses(M, N, int(1.3 * max(nl, nr)), rng)
    lk, lv = _vector(M, N, nl, rng, vl, pool)
    rk, rv = _vector(M, N, nr, rng, vr, pool)
    left = _engine(gpu, kind, M, N, vl, lk, lv)
    right = _engine(gpu, kind, M, N, vr, rk, rv)
    dev, orc = _ops(gpu, kind, M, N)
    got = gpu.observables.overlap_values(
        gpu.DeviceDVec(left, device_model(gpu, kind, M, N, 4.0, 1.0), None),
        gpu.DeviceDVec(right, device_model(gpu, kind, M, N, 4.0, 1.0), None), dev)
    left.close()
    right.close()
    want = O.overlaps(lk, lv, rk, rv, M, orc)
    scale = np.abs(lv).sum() * np.abs(rv).max() * 10
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-13 * scale)
    if vl == "int" and vr == "int":
        assert got[0] == want[0]
    if shared:
        assert got[0] != 0


def test_overlap_edge_cases_and_errors(gpu):
    M = N = 6
    rng = np.random.default_rng(3)
    k, v = _vector(M, N, 200, rng, "float")
    h = device_model(gpu, "real", M, N, 4.0, 1.0)
    a = _engine(gpu, "real", M, N, "float", k, v)
    e = _engine(gpu, "real", M, N, "float", k[:0], v[:0])
    A = gpu.DeviceDVec(a, h, None)
    E = gpu.DeviceDVec(e, h, None)
    # dot with itself, and with an empty vector
    assert gpu.dot(A, A) == pytest.approx(float(np.dot(v, v)), rel=1e-13)
    assert gpu.dot(A, E) == 0.0 and gpu.dot(E, A) == 0.0
    assert gpu.dot(A, h, E) == 0.0
    # the Rayleigh quotient of H from the device equals the oracle's
    want = O.overlaps(k, v, k, v, M, [("ham", oracle_model("real", M, N, 4.0, 1.0))])
    assert gpu.dot(A, h, A) == pytest.approx(want[1], rel=1e-12)
    # a tuple of operators returns a tuple, as dot(t, ops::Tuple, u)
    g = gpu.dot(A, (gpu.G2RealCorrelator(0), gpu.G2RealCorrelator(1)), A)
    assert isinstance(g, tuple) and len(g) == 2
    # more than eight operators are taken in batches
    many = [gpu.DensityMatrixDiagonal(m) for m in range(1, M + 1)] + [gpu.G2RealCorrelator(d) for d in range(M)]
    vals = gpu.observables.overlap_values(A, A, many)
    assert sum(vals[1:M + 1]) == pytest.approx(N * vals[0], rel=1e-12)          # sum_m n_m = N
    assert sum(vals[M + 1:]) == pytest.approx(N * (N - 1) / M * vals[0], rel=1e-12)
    with pytest.raises(ValueError):
        gpu.dot(A, gpu.G2RealCorrelator(M), A)
    with pytest.raises(ValueError):
        gpu.dot(A, gpu.DensityMatrixDiagonal(M + 1), A)
    h7 = device_model(gpu, "real", 7, 6, 4.0, 1.0)
    b = gpu._native.Engine(h7.descriptor(), gpu.IsDeterministic().descriptor(), 1, 10, 10)
    with pytest.raises(gpu.RimuError):
        gpu._native.overlaps(a, b, [])
    b.close()
    a.close()
    e.close()


def _num_stats(df):
    return sum(1 for c in df.columns if c.startswith("c") and c[1:2].isdigit())


def test_all_overlaps_columns(gpu):
    """test/lomc.jl:98-158: report columns c{i}_dot_c{j}, c{i}_Op{k}_c{j}, per-replica
    suffixes; G2RealCorrelator stands in for the Gutzwiller-transformed operator."""
    address = gpu.BoseFS(1, 1, 1, 1, 1, 1, 1)
    H = gpu.HubbardReal1D(address)
    G = gpu.G2RealCorrelator(1)
    dv = gpu.DVec((address, 1), style=gpu.IsDynamicSemistochastic())
    df, sim = gpu.lomc(H, dv, replica_strategy=gpu.NoStats(1), laststep=20)
    sim.close()
    assert "shift" in df.columns and "shift_1" not in df.columns
    df, sim = gpu.lomc(H, dv, replica_strategy=gpu.NoStats(3), laststep=20)
    sim.close()
    assert len(sim.replicas) == 3 and "shift_4" not in df.columns
    assert not df["shift_1"].equals(df["shift_2"]) and not df["shift_2"].equals(df["shift_3"])
    for n, kw, per_pair in [(4, {}, 1), (5, {}, 1), (4, dict(operator=H, vecnorm=False), 1),
                            (4, dict(vecnorm=False), 0), (4, dict(operator=H), 2),
                            (2, dict(operator=(G, H)), 3), (7, dict(operator=[G, H]), 3)]:
        df, sim = gpu.lomc(H, dv, replica_strategy=gpu.AllOverlaps(n, **kw), laststep=20)
        sim.close()
        assert _num_stats(df) == per_pair * math.comb(n, 2), (n, kw)
    # the reported overlaps are those of the replicas' vectors at that step
    df, sim = gpu.lomc(H, dv, replica_strategy=gpu.AllOverlaps(3, operator=H), laststep=15)
    v = sim.state_vectors
    assert df["c1_dot_c3"].iloc[-1] == pytest.approx(gpu.dot(v[0], v[2]), rel=1e-14)
    assert df["c2_Op1_c3"].iloc[-1] == pytest.approx(gpu.dot(v[1], H, v[2]), rel=1e-14)
    sim.close()


def test_replica_g2_estimator(gpu):
    """test/StatsTools.jl:260-290 with h = 0 and more walkers: HubbardReal1D(1,1,1,1),
    u = 6, two replicas with G2RealCorrelator(0) and (1); the replica estimator matches the
    exact ground-state values 0.2207082 and 0.9074473 (the reference quotes 0.220679 and
    0.907466 from a long run)."""
    ham = gpu.HubbardReal1D(gpu.BoseFS(1, 1, 1, 1), u=6.0, t=1.0)
    dv = gpu.DVec((ham.starting_address(), 2))
    skip, run = 2 ** 8, 2 ** 12
    p = gpu.ProjectorMonteCarloProblem(
        ham, start_at=dv, last_step=skip + run, random_seed=179,
        shift_strategy=gpu.DoubleLogUpdate(target_walkers=500),
        replica_strategy=gpu.AllOverlaps(2, operator=(gpu.G2RealCorrelator(0), gpu.G2RealCorrelator(1))))
    sim = gpu.solve(p)
    df = sim.df
    sim.close()
    for op, exact in (("Op1", 0.2207082), ("Op2", 0.9074473)):
        r = ST.rayleigh_replica_estimator(df, op_name=op, skip=skip)
        assert abs(r.pmedian() - exact) < 3 * r.err + 2e-3, (op, r)
