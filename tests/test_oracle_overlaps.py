"""CPU tests that pin the oracle's overlaps (or_overlaps: dot and dot_from_right,
DictVectors/abstractdvec.jl:313-324) and its diagonal observables before the GPU reductions
are checked against it.

Anchors: the brute-force second-quantised Hamiltonians of test_oracle.py (numpy, built from
a+ a on occupation tuples) give dot(t, H, u) as t^T H u; G2RealCorrelator obeys its
normalisation sum_d G2(d) = N(N-1)/M on every address (correlation_functions.jl:49-52) and
its ground-state values in the reference's test (test/StatsTools.jl:260-264, 0.220679 and
0.907466 for HubbardReal1D(1,1,1,1), u = 6, from a long run); DensityMatrixDiagonal sums to
N over the modes.
"""
import itertools

import numpy as np
import pytest

import oracle as O
from test_oracle import brute_mom, brute_real


def _random_pair(B, rng, frac=0.5, ints=False):
    n = len(B)
    out = []
    for _ in range(2):
        idx = np.sort(rng.choice(n, max(1, int(frac * n)), replace=False))
        if ints:
            vals = rng.integers(-9, 10, size=len(idx)).astype(np.int64)
            vals[vals == 0] = 2
        else:
            vals = rng.normal(size=len(idx))
        out.append((idx, vals))
    return out


@pytest.mark.parametrize("kind,M,N", [("real", 5, 4), ("real", 6, 3), ("mom", 5, 4), ("mom", 4, 4)])
@pytest.mark.parametrize("ints", [False, True])
def test_overlaps_match_dense_brute_force(kind, M, N, ints):
    rng = np.random.default_rng(M * 10 + N + ints)
    u, t = 3.0, 1.2
    if kind == "real":
        B, H = brute_real(M, N, u, t)
        model = O.real1d(M, N, u, t)
    else:
        _, kes = O.mom1d_grid(M, t)
        B, H = brute_mom(M, N, u, t, kes)
        model = O.mom1d(M, N, u, t)
    keys = np.array([O.from_onr(b) for b in B], dtype=np.uint64)
    (ti, tv), (ui, uv) = _random_pair(B, rng, 0.6, ints)
    T = np.zeros(len(B))
    U = np.zeros(len(B))
    T[ti] = tv
    U[ui] = uv
    g2 = np.array([[sum(b[i] * (b[(i + d) % M] - (d == 0)) for i in range(M)) / M for d in range(M)]
                   for b in B])
    dens = np.array([[b[m] for m in range(M)] for b in B], dtype=float)
    ops = [("ham", model), ("g2", 0), ("g2", 1), ("density", 2)]
    out = O.overlaps(keys[ti], tv, keys[ui], uv, M, ops)
    want = [T @ U, T @ H @ U, T @ (g2[:, 0] * U), T @ (g2[:, 1] * U), T @ (dens[:, 1] * U)]
    np.testing.assert_allclose(out, want, rtol=1e-12, atol=1e-12)
    if ints:
        assert out[0] == float(int(np.dot(T.astype(np.int64), U.astype(np.int64))))


@pytest.mark.parametrize("M,N", [(4, 4), (6, 3), (5, 7), (9, 2)])
def test_g2_normalisation_and_density_sum_every_address(M, N):
    for b in (o for o in itertools.product(range(N + 1), repeat=M) if sum(o) == N):
        k = O.from_onr(b)
        s = sum(O.g2_real(k, M, d) for d in range(M))
        assert s == pytest.approx(N * (N - 1) / M, rel=1e-14)
        assert sum(O.density_diagonal(k, M, m) for m in range(1, M + 1)) == N
        # periodic: G2(-d) = G2(d) = G2(M - d)
        for d in range(1, M):
            assert O.g2_real(k, M, d) == O.g2_real(k, M, M - d)


def test_g2_ground_state_values_of_reference_test():
    # test/StatsTools.jl:260-264: HubbardReal1D(BoseFS(1,1,1,1)), u = 6, t = 1
    M = N = 4
    B, H = brute_real(M, N, 6.0, 1.0)
    w, V = np.linalg.eigh(H)
    psi = V[:, 0]
    keys = np.array([O.from_onr(b) for b in B], dtype=np.uint64)
    out = O.overlaps(keys, psi, keys, psi, M, [("g2", 0), ("g2", 1), ("ham", O.real1d(M, N, 6.0, 1.0))])
    assert out[0] == pytest.approx(1.0, rel=1e-12)
    # exact 0.2207082 and 0.9074473 against the reference's long-run 0.220679, 0.907466
    assert out[1] == pytest.approx(0.220679, abs=1e-4)
    assert out[2] == pytest.approx(0.907466, abs=1e-4)
    assert out[3] == pytest.approx(w[0], rel=1e-12)
