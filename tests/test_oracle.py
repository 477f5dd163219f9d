"""CPU tests that pin the oracle (oracle/liboracle.so + oracle/*.py) before it is trusted.

Every check is anchored on something the oracle's own formulas do not restate: the
reference's doctests and test fixtures (cited), published known-answer vectors, exact
diagonalisation energies held by the reference's tests, and a brute-force second-quantised
Hamiltonian built independently with numpy.
"""
import itertools
import math

import numpy as np
import pytest

import oracle as O
from rimu_amd import statstools as ST


# ----------------------------------------------------------------------------------------
# Philox4x32-10: known-answer vectors of the Random123 distribution (kat_vectors,
# philox4x32_10 rows), the published reference of Salmon et al. (SC'11).
# ----------------------------------------------------------------------------------------
@pytest.mark.parametrize("ctr,key,out", [
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff, 0xffffffff), (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
])
def test_philox_known_answers(ctr, key, out):
    assert O.philox4x32_10(ctr, key) == out


# ----------------------------------------------------------------------------------------
# BoseFS encoding
# ----------------------------------------------------------------------------------------
def test_bosefs_layout_and_occupied_modes():
    # fockaddress.jl doctest: BoseFS{13,7}(10,0,0,0,2,0,1) -> (10,1,0), (2,5,14), (1,7,18)
    b = O.from_onr((10, 0, 0, 0, 2, 0, 1))
    assert O.occupied_modes(b) == [(10, 1, 0), (2, 5, 14), (1, 7, 18)]
    assert O.to_onr(b, 7) == (10, 0, 0, 0, 2, 0, 1)
    # test/BitStringAddresses.jl:220-229 shape: mode 1 in the low bits, separators are zeros
    assert O.from_onr((0, 1, 2, 3, 0)) == 0b111011010
    # bose_hubbard_interaction doctests (bosefs.jl:393-398)
    assert O.bose_hubbard_interaction(O.from_onr((2, 1, 1, 0))) == 2
    assert O.bose_hubbard_interaction(O.from_onr((3, 0, 1, 0))) == 6
    # test/BitStringAddresses.jl:202-206 (single-chunk analogues of the fixtures)
    assert O.bose_hubbard_interaction(O.from_onr((0, 8, 0, 0, 2))) == 8 * 7 + 2


def test_bosefs_round_trip_all_small_addresses():
    for M, N in [(1, 5), (5, 1), (4, 6), (7, 3)]:
        for onr in itertools.product(range(N + 1), repeat=M):
            if sum(onr) != N:
                continue
            b = O.from_onr(onr)
            assert bin(b).count("1") == N and b < 2 ** (N + M - 1)
            assert O.to_onr(b, M) == onr
            occ = O.occupied_modes(b)
            assert [(n, m) for n, m, _ in occ] == [(n, i + 1) for i, n in enumerate(onr) if n]


# ----------------------------------------------------------------------------------------
# Hamiltonian doctests
# ----------------------------------------------------------------------------------------
def test_hopnextneighbour_doctests():
    # bosefs.jl:281-285: hopnextneighbour(BoseFS(1,0,1), 3) == ((2,0,0), sqrt 2); 4 -> ((1,1,0), 1)
    h = O.real1d(3, 2, 1.0, 1.0)
    a = O.from_onr((1, 0, 1))
    nb, v = O.get_offdiagonal(h, a, 3)
    assert O.to_onr(nb, 3) == (2, 0, 0) and v == -math.sqrt(2)
    nb, v = O.get_offdiagonal(h, a, 4)
    assert O.to_onr(nb, 3) == (1, 1, 0) and v == -1.0


def test_hubbardreal1d_offdiagonals_doctest():
    # Interfaces/hamiltonians.jl:323-330: offdiagonals(HubbardReal1D(BoseFS(3,2,1)))
    expect = [((2, 3, 1), -3.0), ((2, 2, 2), -2.449489742783178), ((3, 1, 2), -2.0),
              ((4, 1, 1), -2.8284271247461903), ((4, 2, 0), -2.0), ((3, 3, 0), -1.7320508075688772)]
    h = O.real1d(3, 6, 1.0, 1.0)
    got = [(O.to_onr(b, 3), v) for b, v in O.offdiagonals(h, O.from_onr((3, 2, 1)))]
    assert got == expect


def test_hubbardmom1d_doctests():
    h = O.mom1d(3, 6, 1.0, 1.0)
    a = O.from_onr((3, 2, 1))
    assert O.diagonal_element(h, a) == 8.666666666666664      # Interfaces/hamiltonians.jl:236
    assert O.num_offdiagonals(h, a) == 10                       # :257
    nb, v = O.get_offdiagonal(h, a, 3)                          # :278
    assert O.to_onr(nb, 3) == (2, 1, 3) and v == 1.0
    # momentum_transfer_diagonal(H, map) for BoseFS{6,5}(1,2,3,0,0) == 5.2 (HubbardMom1D.jl:154-161)
    h5 = O.mom1d(5, 6, 1.0, 1.0)
    b = O.from_onr((1, 2, 3, 0, 0))
    kin = sum(h5.kes[j] * n for j, n in enumerate((1, 2, 3, 0, 0)))
    assert O.diagonal_element(h5, b) - kin == pytest.approx(5.2, rel=1e-15)
    # num_singly_doubly_occupied_sites doctests (HubbardMom1D.jl:96-101) through n_od
    assert O.num_offdiagonals(O.mom1d(3, 3), O.from_onr((1, 1, 1))) == 3 * 2 * 1
    assert O.num_offdiagonals(O.mom1d(3, 3), O.from_onr((2, 0, 1))) == 2 * 1 * 1 + 1 * 2


def test_momentum_transfer_excitation_fixtures():
    # test/Hamiltonians.jl:98-122 (fold=true); u = 2M makes the value the bare onproduct
    M = 4
    h = O.mom1d(M, 2, u=2 * M)
    a1 = O.from_onr((0, 1, 1, 0))
    for i in range(1, 5):
        nb, v = O.get_offdiagonal(h, a1, i)
        assert O.to_onr(nb, 4) == (1, 0, 0, 1) and v == 1.0
    a3 = O.from_onr((1, 1, 0, 0))
    for i in range(1, 5):
        assert O.get_offdiagonal(h, a3, i)[1] == 1.0
    h3 = O.mom1d(3, 3, u=6.0)
    a4 = O.from_onr((0, 3, 0))
    for i in (1, 2):
        nb, v = O.get_offdiagonal(h3, a4, i)
        assert O.to_onr(nb, 3) == (1, 1, 1) and v == pytest.approx(math.sqrt(6), rel=1e-15)


def test_t_u_zero_limits():
    # test/Hamiltonians.jl:217-228
    assert O.diagonal_element(O.mom1d(3, 1, t=0.0), O.from_onr((0, 1, 0))) == 0
    assert O.diagonal_element(O.real1d(3, 9, u=0.0), O.from_onr((3, 3, 3))) == 0
    assert all(v == 0 for _, v in O.offdiagonals(O.mom1d(3, 1, u=0.0), O.from_onr((0, 1, 0))))
    assert all(v == 0 for _, v in O.offdiagonals(O.real1d(3, 9, t=0.0), O.from_onr((3, 3, 3))))
    h = O.mom1d(3, 3, u=0.0, t=1.0)
    assert 2 * 1.0 * 3 + O.diagonal_element(h, O.from_onr((0, 3, 0))) == pytest.approx(0, abs=1e-15)


# ----------------------------------------------------------------------------------------
# Independent brute force: second-quantised H built from the textbook definitions
# ----------------------------------------------------------------------------------------
def _basis(M, N):
    return [o for o in itertools.product(range(N + 1), repeat=M) if sum(o) == N]


def _annihilate(state, i):
    n = state[i]
    if n == 0:
        return None, 0.0
    s = list(state)
    s[i] -= 1
    return tuple(s), math.sqrt(n)


def _create(state, i):
    s = list(state)
    s[i] += 1
    return tuple(s), math.sqrt(s[i])


def brute_real(M, N, u, t):
    """H = -t sum_<ij> a+_i a_j (periodic chain) + u/2 sum n(n-1)."""
    B = _basis(M, N)
    idx = {b: k for k, b in enumerate(B)}
    H = np.zeros((len(B), len(B)))
    for b in B:
        c = idx[b]
        H[c, c] += u / 2 * sum(n * (n - 1) for n in b)
        for i in range(M):
            for j in ((i + 1) % M, (i - 1) % M):
                s1, f1 = _annihilate(b, i)
                if s1 is None:
                    continue
                s2, f2 = _create(s1, j)
                H[idx[s2], c] += -t * f1 * f2
    return B, H


def brute_mom(M, N, u, t, kes):
    """H = sum_k e_k n_k + u/(2M) sum_{p,q,k} a+_{p+k} a+_{q-k} a_q a_p (modes mod M)."""
    B = _basis(M, N)
    idx = {b: k for k, b in enumerate(B)}
    H = np.zeros((len(B), len(B)))
    for b in B:
        c = idx[b]
        H[c, c] += sum(kes[i] * n for i, n in enumerate(b))
        for p in range(M):
            for q in range(M):
                for k in range(M):
                    s1, f1 = _annihilate(b, p)
                    if s1 is None:
                        continue
                    s2, f2 = _annihilate(s1, q)
                    if s2 is None:
                        continue
                    s3, f3 = _create(s2, (q - k) % M)
                    s4, f4 = _create(s3, (p + k) % M)
                    H[idx[s4], c] += u / (2 * M) * f1 * f2 * f3 * f4
    return B, H


def oracle_matrix(model, B):
    M = len(B[0])
    keys = [O.from_onr(b) for b in B]
    idx = {k: i for i, k in enumerate(keys)}
    H = np.zeros((len(B), len(B)))
    for c, k in enumerate(keys):
        H[c, c] += O.diagonal_element(model, k)
        for nb, v in O.offdiagonals(model, k):
            H[idx[nb], c] += v
    return H


@pytest.mark.parametrize("M,N,u,t", [(4, 3, 1.7, 0.9), (5, 5, 6.0, 1.0), (2, 3, 1.0, 1.0), (6, 2, 0.3, 2.0)])
def test_real1d_matches_brute_force(M, N, u, t):
    B, Hb = brute_real(M, N, u, t)
    Ho = oracle_matrix(O.real1d(M, N, u, t), B)
    np.testing.assert_allclose(Ho, Hb, rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize("M,N,u,t", [(5, 4, 2.3, 1.1), (6, 3, 4.0, 1.0), (4, 4, 1.0, 0.5), (3, 5, 6.0, 1.0)])
def test_mom1d_matches_brute_force(M, N, u, t):
    _, kes = O.mom1d_grid(M, t)
    B, Hb = brute_mom(M, N, u, t, kes)
    Ho = oracle_matrix(O.mom1d(M, N, u, t), B)
    np.testing.assert_allclose(Ho, Hb, rtol=1e-13, atol=1e-13)


def test_mom1d_grid_is_julia_range():
    # k grid (HubbardMom1D.jl:55-62): even M puts k = 0 at mode M/2, odd M at (M+1)/2
    ks, kes = O.mom1d_grid(10, 1.0)
    assert ks[4] == 0.0 and ks[-1] == pytest.approx(math.pi, abs=1e-15)
    ks5, _ = O.mom1d_grid(5, 1.0)
    assert ks5[2] == pytest.approx(0.0, abs=1e-15)
    np.testing.assert_allclose(kes, -2 * np.cos(ks), rtol=0, atol=1e-15)


# ----------------------------------------------------------------------------------------
# Exact diagonalisation against energies held by the reference's tests
# ----------------------------------------------------------------------------------------
def sector_basis(M, N, K):
    """Momentum sector sum_j j*n_j = K (mod M) of the full basis."""
    basis = O.build_basis(M, N)
    keep = [b for b in basis.tolist() if sum((j + 1) * n for j, n in enumerate(O.to_onr(b, M))) % M == K % M]
    return np.array(keep, dtype=np.uint64)


@pytest.mark.parametrize("M,N,u,E0,tol", [
    (5, 5, 1.0, -8.280991746582686, 1e-11),   # test/lomc.jl:546
    (7, 7, 6.0, -4.628524493494574, 1e-11),   # test/mpi_runtests.jl:160
    (6, 6, 6.0, -4.0215, 1e-4),               # test/KrylovKit.jl:15
])
def test_ed_real1d_known_answers(M, N, u, E0, tol):
    model = O.real1d(M, N, u, 1.0)
    e = O.ed_energies(model, O.build_basis(M, N), 1)[0]
    assert e == pytest.approx(E0, abs=tol)


def test_ed_real1d_spectrum_doctest():
    # exact_diagonalization_problem.jl:86: HubbardReal1D(BoseFS(1,1,1)) spectrum
    expect = [-5.09593, -1.51882, -1.51882, 1.55611, 1.6093, 1.6093, 4.0, 4.53982, 4.90952, 4.90952]
    e = O.ed_energies(O.real1d(3, 3), O.build_basis(3, 3), 10)
    np.testing.assert_allclose(e, expect, atol=6e-6)


def test_ed_mom1d_known_answers():
    # test/ExactDiagonalization.jl:189: HubbardMom1D(BoseFS(1,2,3)) -> -3.045633163020568
    K = sum((j + 1) * n for j, n in enumerate((1, 2, 3)))
    e = O.ed_energies(O.mom1d(3, 6), sector_basis(3, 6, K), 1)[0]
    assert e == pytest.approx(-3.045633163020568, abs=1e-11)
    # pdvec.jl:136-141: HubbardMom1D(BoseFS(0,0,5,0,0); u=6) lowest four
    K = 3 * 5
    e = O.ed_energies(O.mom1d(5, 5, 6.0), sector_basis(5, 5, K), 4)
    np.testing.assert_allclose(e, [-3.4311156892322234, 1.1821748602612363, 3.7377753753082823,
                                   6.996390417443125], atol=1e-11)


@pytest.mark.parametrize("u,E0", [(4.0, -9.251592973178997), (1.0, -16.36048582876015)])
def test_ed_mom1d_10x10_known_answers(u, E0):
    # test/lomc.jl:643 and :679: condensate BoseFS{10,10}(0,0,0,0,10,0,...)
    e = O.ed_energies(O.mom1d(10, 10, u), sector_basis(10, 10, 5 * 10), 1)[0]
    assert e == pytest.approx(E0, abs=1e-9)


def test_bfs_basis_equals_momentum_sector():
    model = O.mom1d(6, 4, 2.0)
    start = O.from_onr((0, 0, 4, 0, 0, 0))
    assert np.array_equal(O.bfs_basis(model, start), sector_basis(6, 4, 3 * 4))


# ----------------------------------------------------------------------------------------
# Step arithmetic: closed forms of diagonal_step! (test/StochasticStyles.jl:128-153)
# ----------------------------------------------------------------------------------------
def _single_step(model, sty, onr, val, shift, dtau, seed=1):
    k = O.from_onr(onr)
    return O.step(model, sty, [k], [val], seed=seed, step=1, shift=shift, dtau=dtau)


@pytest.mark.parametrize("shift,dtau,val,out,stat,stat_val", [
    (10.0, 1.0, 2.5, 25.0, "clones", 22.5),
    (-0.5, 0.5, 1.0, 0.25, "deaths", 0.75),
    (-10.0, 0.5, 1.0, -4.5, "zombies", 4.5),
])
def test_diagonal_step_exact(shift, dtau, val, out, stat, stat_val):
    model = O.real1d(3, 3, 1.0, 0.0)  # t = 0: only the diagonal acts; H_aa(2,0,1) = 1
    k, v, st = _single_step(model, O.style("deterministic"), (2, 0, 1), val, shift, dtau)
    assert list(v) == [out]
    assert st[stat] == stat_val


def test_diagonal_step_integer_rounding():
    model = O.real1d(3, 3, 1.0, 0.0)
    seen = set()
    for seed in range(40):
        k, v, st = _single_step(model, O.style("integer"), (2, 0, 1), 1, 10.0, 0.5, seed=seed)
        assert st["clones"] in (4, 5) and v[0] == st["clones"] + 1   # x = 5.5
        seen.add(int(v[0]))
    assert seen == {5, 6}
    for seed in range(40):  # zombies: x = -4.5 -> -4 or -5
        k, v, st = _single_step(model, O.style("integer"), (2, 0, 1), 1, -10.0, 0.5, seed=seed)
        assert st["deaths"] == 1 and st["zombies"] in (4, 5) and v[0] == 1 - 1 - st["zombies"]


# ----------------------------------------------------------------------------------------
# Stochastic pieces are unbiased (test/StochasticStyles.jl:187-275)
# ----------------------------------------------------------------------------------------
@pytest.mark.parametrize("model,onr", [
    (O.mom1d(6, 5, 6.0), (0, 0, 0, 3, 1, 1)),
    (O.real1d(7, 10, 1.0), (5, 0, 0, 5, 0, 0, 0)),
])
@pytest.mark.parametrize("kind", ["integer", "dynamic"])
def test_spawning_unbiased_against_exact(model, onr, kind):
    k0 = O.from_onr(onr)
    val = 7 if kind == "integer" else 7.0
    dk, dv, _ = O.step(model, O.style("deterministic"), [k0], [float(val)], seed=0, step=1,
                       shift=0.5, dtau=0.01)
    exact = dict(zip(dk.tolist(), dv.tolist()))
    acc = {}
    n = 4000
    sty = O.style(kind, compression=0) if kind == "dynamic" else O.style(kind)
    for seed in range(n):
        k, v, _ = O.step(model, sty, [k0], [val], seed=seed, step=1, shift=0.5, dtau=0.01)
        for kk, vv in zip(k.tolist(), v.tolist()):
            acc[kk] = acc.get(kk, 0.0) + vv
    total_exact = sum(abs(x) for x in exact.values())
    for kk, ev in exact.items():
        assert acc.get(kk, 0.0) / n == pytest.approx(ev, rel=0.1, abs=0.02 * total_exact)


def test_projected_deposits_unbiased():
    """IsDynamicSemistochastic(late_compression=false) projects every deposit below the
    threshold to +-threshold with probability |x|/threshold (projected_deposit!,
    spawning.jl:13-31), so the mean over seeds equals the exact product.  With a huge
    rel_spawning_threshold every column is exact, so the projection is the only noise."""
    model = O.real1d(6, 6, 2.0, 1.0)
    rng = np.random.default_rng(3)
    basis = O.build_basis(6, 6)
    keys = np.sort(rng.choice(basis, 60, replace=False))
    vals = rng.uniform(-2.0, 2.0, size=60)
    dk, dv, _ = O.step(model, O.style("deterministic"), keys, vals, seed=0, step=1,
                       shift=0.3, dtau=0.05)
    exact = dict(zip(dk.tolist(), dv.tolist()))
    sty = O.style("dynamic", proj_threshold=1.0, rel_threshold=1e9, compression=0)
    n = 2000
    acc, acc2 = {}, {}
    for seed in range(n):
        k, v, st = O.step(model, sty, keys, vals, seed=seed, step=1, shift=0.3, dtau=0.05)
        assert st["exact_steps"] == 60 and st["inexact_steps"] == 0
        for kk, vv in zip(k.tolist(), v.tolist()):
            acc[kk] = acc.get(kk, 0.0) + vv
            acc2[kk] = acc2.get(kk, 0.0) + vv * vv
    for kk, ev in exact.items():
        mean = acc.get(kk, 0.0) / n
        var = max(acc2.get(kk, 0.0) / n - mean * mean, 1e-12)
        assert abs(mean - ev) < 5 * math.sqrt(var / n) + 1e-12, kk


@pytest.mark.parametrize("val,rel,absth,boost,exact", [
    # n_od = 12 for the six singly occupied modes of near_uniform(6, 6)
    (12.0, 1.0, math.inf, 1.0, True),    # amount 12 >= min(Inf, 12)
    (11.5, 1.0, math.inf, 1.0, False),
    (6.5, 2.0, math.inf, 1.0, True),     # rel_threshold multiplies the walker number
    (5.5, 2.0, math.inf, 1.0, False),
    (3.5, 1.0, 3.0, 1.0, True),          # abs_threshold caps the threshold
    (2.5, 1.0, 3.0, 1.0, False),
    (6.5, 1.0, math.inf, 2.0, True),     # boost multiplies the amount (spawning.jl:366)
    (1.6, 1.0, 3.0, 2.0, True),          # ... but not abs_threshold
    (1.4, 1.0, 3.0, 2.0, False),
])
def test_dynamic_semistochastic_exact_rule(val, rel, absth, boost, exact):
    """spawn!(DynamicSemistochastic) (spawning.jl:364-378) spawns exactly when
    boost * |v| * rel_threshold >= min(abs_threshold, num_offdiagonals); the docstrings of
    IsDynamicSemistochastic (styles.jl:158-164) and DynamicSemistochastic
    (spawning.jl:343-349) state which thresholds boost scales.  An exact column equals the
    deterministic product; a stochastic one draws max(floor(boost |v|), 1) attempts."""
    model = O.real1d(6, 6, 2.0, 1.0)
    k0 = O.from_onr((1, 1, 1, 1, 1, 1))
    assert O.num_offdiagonals(model, k0) == 12
    sty = O.style("dynamic", rel_threshold=rel, abs_threshold=absth, compression=0)
    k, v, st = O.step(model, sty, [k0], [val], seed=9, step=1, shift=0.0, dtau=0.01, boost=boost)
    dk, dv, _ = O.step(model, O.style("deterministic"), [k0], [val], seed=0, step=1,
                       shift=0.0, dtau=0.01)
    assert st["exact_steps"] == (1 if exact else 0)
    assert st["inexact_steps"] == (0 if exact else 1)
    if exact:
        assert st["attempts"] == 12
        assert np.array_equal(k, dk) and np.array_equal(v, dv)
    else:
        assert st["attempts"] == max(math.floor(boost * val), 1)


def test_boost_unbiased_integer():
    """boost multiplies the number of attempts of IsStochasticInteger
    (spawn!(WithReplacement), spawning.jl:232-243) without changing the expected spawn."""
    model = O.real1d(7, 10, 1.0)
    k0 = O.from_onr((5, 0, 0, 5, 0, 0, 0))
    dk, dv, _ = O.step(model, O.style("deterministic"), [k0], [7.0], seed=0, step=1,
                       shift=0.5, dtau=0.01)
    exact = dict(zip(dk.tolist(), dv.tolist()))
    acc = {}
    n = 3000
    for seed in range(n):
        k, v, st = O.step(model, O.style("integer"), [k0], [7], seed=seed, step=1, shift=0.5,
                          dtau=0.01, boost=2.0)
        assert st["attempts"] == 14
        for kk, vv in zip(k.tolist(), v.tolist()):
            acc[kk] = acc.get(kk, 0.0) + vv
    total_exact = sum(abs(x) for x in exact.values())
    for kk, ev in exact.items():
        assert acc.get(kk, 0.0) / n == pytest.approx(ev, rel=0.1, abs=0.02 * total_exact)


def test_threshold_compression_keeps_mean():
    model = O.real1d(6, 6, 1.0, 1.0)
    rng = np.random.default_rng(0)
    basis = O.build_basis(6, 6)
    keys = np.sort(rng.choice(basis, 200, replace=False))
    vals = rng.uniform(-1.5, 1.5, size=200)
    dk, dv, _ = O.step(model, O.style("dynamic", compression=0), keys, vals, seed=0, step=1,
                       shift=0.0, dtau=0.01)
    ref = dict(zip(dk.tolist(), dv.tolist()))
    acc = {}
    n = 1000
    for seed in range(n):
        k, v, st = O.step(model, O.style("dynamic"), keys, vals, seed=seed, step=1, shift=0.0, dtau=0.01)
        assert st["len"] <= st["len_before"]
        assert all(abs(x) >= 1.0 for x in v)
        for kk, vv in zip(k.tolist(), v.tolist()):
            acc[kk] = acc.get(kk, 0.0) + vv
    # Dynamic spawning is itself stochastic below n_od walkers; compare the grand mean
    assert sum(acc.values()) / n == pytest.approx(sum(ref.values()), rel=0.05)


# ----------------------------------------------------------------------------------------
# Orders of deposits and the threaded restatement
# ----------------------------------------------------------------------------------------
def test_column_order_and_list_order_agree():
    model = O.real1d(8, 8, 3.0, 1.0)
    rng = np.random.default_rng(1)
    keys = np.sort(rng.choice(O.build_basis(8, 8), 500, replace=False))
    vi = rng.integers(-9, 10, size=500)
    vi[vi == 0] = 2
    a = O.step(model, O.style("integer"), keys, vi, seed=3, step=2, shift=-1.0, dtau=0.01, order=0)
    b = O.step(model, O.style("integer"), keys, vi, seed=3, step=2, shift=-1.0, dtau=0.01, order=1)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])   # integers are exact
    vf = rng.normal(size=500) * 5
    a = O.step(model, O.style("deterministic"), keys, vf, seed=3, step=2, shift=-1.0, dtau=0.01, order=0)
    b = O.step(model, O.style("deterministic"), keys, vf, seed=3, step=2, shift=-1.0, dtau=0.01, order=1)
    assert np.array_equal(a[0], b[0])
    np.testing.assert_allclose(a[1], b[1], rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("kind", ["integer", "dynamic"])
def test_threaded_restatement_equals_serial(kind):
    model = O.real1d(10, 10, 6.0, 1.0)
    rng = np.random.default_rng(2)
    keys = np.sort(rng.choice(O.build_basis(10, 10), 3000, replace=False))
    vals = rng.integers(1, 6, size=3000) if kind == "integer" else rng.exponential(3.0, size=3000)
    s = O.step(model, O.style(kind), keys, vals, seed=9, step=4, shift=-2.0, dtau=0.01, order=0)
    t = O.step(model, O.style(kind), keys, vals, seed=9, step=4, shift=-2.0, dtau=0.01, threads=4)
    order = np.argsort(t[0])
    assert np.array_equal(s[0], t[0][order])
    if kind == "integer":
        assert np.array_equal(s[1], t[1][order])
    else:
        np.testing.assert_allclose(s[1], t[1][order], rtol=1e-12)
    assert s[2]["attempts"] == t[2]["attempts"] and s[2]["len"] == t[2]["len"]


def test_annihilation_fold_semantics():
    # deposit! then delete-on-zero (dvec.jl:129-136): cancellations vanish, order kept
    k = np.array([5, 3, 5, 3, 9, 5], dtype=np.uint64)
    v = np.array([1.5, 2.0, -1.5, 0.25, 0.0, 1e-300])
    ok, ov = O.annihilate(k, v)
    assert ok.tolist() == [3, 5] and ov.tolist() == [2.25, 1e-300]
    ki = np.array([1, 1, 2, 2, 2], dtype=np.uint64)
    vi = np.array([3, -3, 1, 1, -5], dtype=np.int64)
    ok, ov = O.annihilate(ki, vi)
    assert ok.tolist() == [2] and ov.tolist() == [-3]


def test_double_log_update_closed_form():
    # shiftstrategy.jl:174-181
    s = O.double_log_update(1.0, 200.0, 100.0, 400.0, zeta=0.08, xi=0.0016, dtau=0.01)
    assert s == pytest.approx(1.0 - (0.0016 / 0.01 * math.log(0.5) + 0.08 / 0.01 * math.log(2.0)))


# ----------------------------------------------------------------------------------------
# Blocking analysis (test/StatsTools.jl:33-59)
# ----------------------------------------------------------------------------------------
def test_blocking_white_and_correlated_noise():
    rng = np.random.default_rng(13)
    v = rng.normal(size=2 ** 10)
    br = ST.blocking_analysis(v)
    assert 0.01 < br.err < 0.04 and br.k <= 4
    brs = ST.blocking_analysis(ST.smoothen(v, 2 ** 5))
    assert brs.mean == pytest.approx(br.mean) and 0.01 < brs.err < 0.04
    assert 5 < brs.k <= 7
    bor = ST.blocking_analysis(np.ones(2000))
    assert bor.k == -1 and math.isnan(bor.err)
    assert ST.blocking_analysis([1.0]).k == -1


def test_ratio_of_means_reference_cases():
    """The set-ups of test/StatsTools.jl:65-153 with numpy draws: zero variance gives
    (0, 0, 0); a ratio of two zero-mean series is Cauchy-like and its 95% interval holds 0;
    a well-behaved correlated pair reproduces x_by_y_linear within 8/sqrt(n); blocking of a
    smoothed pair finds k in 6..8 and keeps the quantiles."""
    r = ST.ratio_of_means([0, 0, 0, 0, 0, 0], [1, 1, 1, 1, 1, 2])
    assert ST.val_and_errs(r) == (0.0, 0.0, 0.0)
    rng = np.random.default_rng(17)
    r = ST.ratio_of_means(rng.normal(size=2000), rng.normal(size=2000))
    q95 = r.pquantile([0.025, 0.975])
    assert q95[0] < 0 < q95[1]
    # Measurements.jl's uncorrelated propagation: |f| sqrt((sa/ma)^2 + (sb/mb)^2)
    f, sf = ST.x_by_y_linear(2.0, 3.0, 0.5, 0.1, 0.0)
    assert f == 2.0 / 3.0
    assert sf == pytest.approx(abs(f) * math.hypot(0.5 / 2.0, 0.1 / 3.0), rel=1e-14)
    n = 2000
    mu_a, mu_b, s_a, s_b, rho = 2.0, 3.0, 0.5, 0.1, 0.02
    f, sf = ST.x_by_y_linear(mu_a, mu_b, s_a, s_b, rho)
    rng = np.random.default_rng(1)  # a fixed seed, as the reference fixes its own
    ab = rng.multivariate_normal([mu_a, mu_b], np.array([[s_a ** 2, rho], [rho, s_b ** 2]]) * n, size=n)
    a, b = ab[:, 0], ab[:, 1]
    r = ST.ratio_of_means(a, b)
    assert 1 <= r.k <= 4 and r.success
    assert r.f == a.mean() / b.mean()
    assert abs(mu_a / mu_b - r.f) < 2 * sf
    assert abs(mu_a / mu_b - r.pmedian()) < 2 * sf
    q = r.pquantile([0.16, 0.5, 0.84])
    assert q[2] - q[1] == pytest.approx(sf, rel=8 / math.sqrt(n))
    assert q[1] - q[0] == pytest.approx(sf, rel=8 / math.sqrt(n))
    assert r.sigma_f == pytest.approx(sf, rel=8 / math.sqrt(n))
    assert abs(r.delta_y) < 0.1
    rs = ST.ratio_of_means(ST.smoothen(a, 2 ** 3), ST.smoothen(b, 2 ** 6))
    assert 6 <= rs.k <= 8
    np.testing.assert_allclose(rs.pquantile([0.16, 0.5, 0.84]), q, atol=2 * sf)
    assert abs(rs.sigma_f - sf) < 2 * sf


def test_ratio_particles_follow_the_ratio_of_normals():
    """The Monte Carlo ratio of a correlated normal pair against its exact distribution
    (Hinkley 1969: the CDF of x/y for bivariate normal (x, y), by numerical integration of
    P(x - z y <= 0) over y), at a noisy denominator (delta_y = 0.25) where linear
    propagation fails.  A dropped covariance term or swapped variances moves the quantiles
    by several times the tolerance."""
    from scipy import integrate, stats
    mx, my, sx, sy, c = 1.0, 2.0, 0.3, 0.5, 0.06
    rng = np.random.default_rng(5)
    xy = ST._particles_2d([mx, my], [[sx ** 2, c], [c, sy ** 2]], 400_000, rng)
    z = ST.Particles(xy[0] / xy[1])
    rho = c / (sx * sy)

    def cdf(t):
        # P(x/y <= t) = int over y of p(y) P(x <= t y | y) for y > 0, P(x >= t y | y) for y < 0
        def integrand(y):
            m = mx + rho * sx / sy * (y - my)
            s = sx * math.sqrt(1 - rho ** 2)
            p = stats.norm.cdf(t * y, loc=m, scale=s)
            return stats.norm.pdf(y, loc=my, scale=sy) * (p if y > 0 else 1 - p)
        return (integrate.quad(integrand, my - 12 * sy, 0, limit=200)[0]
                + integrate.quad(integrand, 0, my + 12 * sy, limit=200)[0])

    for q in (0.025, 0.16, 0.5, 0.84, 0.975):
        assert cdf(float(z.pquantile(q))) == pytest.approx(q, abs=0.004), q
    # delta_y = standard error of the denominator's mean over that mean
    _, _, _, dy, _ = ST.ratio_estimators(np.array([mx + sx, mx - sx]), np.array([my + sy, my - sy]))
    assert dy == pytest.approx(sy / my, rel=1e-12)


def test_val_and_errs_of_a_blocking_result():
    """convenience.jl:61-112: a BlockingResult gives (mean, n*err, n*err)."""
    br = ST.BlockingResult(mean=1.5, err=0.25, err_err=0.0, p_cov=0.0, k=2, blocks=100)
    v, lo, hi = ST.val_and_errs(br, n=2)
    assert v == 1.5 and lo == pytest.approx(0.5, rel=1e-12) and hi == pytest.approx(0.5, rel=1e-12)
    assert ST.within(br, 1.99) and not ST.within(br, 2.01)


def test_mtest_quantiles_match_chi2_table():
    # 99% upper quantiles of chi^2(k): 6.635 (k=1), 9.210 (k=2), 11.345 (k=3)
    from scipy.stats import chi2
    assert chi2.isf(0.01, 1) == pytest.approx(6.6349, abs=1e-4)
    assert chi2.isf(0.01, 2) == pytest.approx(9.2103, abs=1e-4)
    assert chi2.isf(0.01, 3) == pytest.approx(11.3449, abs=1e-4)
