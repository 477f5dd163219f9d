"""CPU tests that pin the oracle's initiator step (or_step_initiator) before the GPU path is
checked against it.

The anchors are independent of the oracle's per-deposit restatement:

* a dense-matrix statement of the rules for IsDeterministic (no randomness): the columns of
  T = 1 + dt (S - H) are split by the parent's initiator status, which fixes the three
  components (safe, unsafe, initiator) of every address as matrix-vector products;
* limits that reduce to the plain step: a threshold no value exceeds leaves only the
  diagonal deposits (the same step with t = 0), a threshold every value exceeds gives the
  plain step exactly for integer walkers;
* the statistical test of the reference, test/lomc.jl:612-675 ("Energies below the plateau
  & initiator bias"): below the plateau the plain run's shift falls under the exact energy
  E0 = -9.251592973178997, every initiator rule biases it above E0, SimpleInitiator the most,
  Initiator and CoherentInitiator alike.
"""
import numpy as np
import pytest

import oracle as O
from rimu_amd import statstools as ST

RULES = ("initiator", "simple", "coherent")


def _dense_T(model, basis, shift, dtau):
    H = O.sparse_hamiltonian(model, basis).toarray()
    return np.eye(len(basis)) + dtau * (shift * np.eye(len(basis)) - H)


def _rule_values(rule, thr, I, S, U):
    if rule == "initiator":
        return S + I + np.where(I != 0, U, 0.0)
    if rule == "simple":
        return S + I
    keep_u = (I != 0) | (np.abs(U) > thr)
    return I + S + np.where(keep_u, U, 0.0)


@pytest.mark.parametrize("rule", RULES)
@pytest.mark.parametrize("kind", ["real", "mom"])
def test_deterministic_step_matches_dense_matrix_split(rule, kind):
    if kind == "real":
        M, N = 6, 4
        model = O.real1d(M, N, 2.0, 1.0)
    else:
        M, N = 6, 4
        model = O.mom1d(M, N, 3.0, 1.0)
    basis = O.build_basis(M, N)
    D = len(basis)
    rng = np.random.default_rng(7 if kind == "real" else 8)
    thr = 1.0
    # values on both sides of the threshold and exactly on it (|v| = thr is not an initiator)
    v = np.zeros(D)
    support = rng.choice(D, size=D // 2 if kind == "real" else D // 8, replace=False)
    mag = rng.uniform(0.2, 4.0, size=len(support))
    mag[::5] = 1.0
    v[support] = mag * rng.choice([-1.0, 1.0], size=len(support))
    shift, dtau = -1.0, 1.0
    T = _dense_T(model, basis, shift, dtau)
    P = np.abs(v) > thr
    Dg = np.diag(T) * v
    Off = T - np.diag(np.diag(T))
    I = np.where(P, Dg, 0.0)
    S = np.where(P, 0.0, Dg) + Off @ np.where(P, v, 0.0)
    U = Off @ np.where(P, 0.0, v)
    exists = (I != 0) | (S != 0) | (U != 0)
    nz = np.concatenate([I[I != 0], S[S != 0], U[U != 0]])
    assert np.all(np.abs(nz) > 1e-9)  # no accidental cancellation in this set-up
    want = _rule_values(rule, thr, I, S, U)
    keys = basis[support]
    order = np.argsort(keys)
    k, val, st = O.step_initiator(model, O.style("deterministic"), keys[order], v[support][order],
                                  rule=rule, threshold=thr, seed=1, step=1, shift=shift,
                                  dtau=dtau, order=0)
    assert k.tolist() == basis[exists].tolist()
    np.testing.assert_allclose(val, want[exists], rtol=1e-12, atol=1e-13)
    assert st["len_before"] == exists.sum() and st["len"] == exists.sum()
    # the set-up exercises every branch: discarded unsafe parts (explicit zeros stored by
    # NoCompression) and, for the coherent rule, unsafe sums on both sides of the threshold
    lone = exists & (I == 0) & (S == 0)
    assert lone.sum() > 0
    if rule != "coherent":
        assert np.all(val[lone[exists]] == 0)
    else:
        assert (lone & (np.abs(U) > thr)).sum() > 0 and (lone & (np.abs(U) <= thr)).sum() > 0


@pytest.mark.parametrize("rule", RULES)
def test_integer_threshold_limits(rule):
    M = N = 8
    model = O.real1d(M, N, 3.0, 1.0)
    rng = np.random.default_rng(11)
    keys = np.sort(rng.choice(O.build_basis(M, N), 400, replace=False))
    vals = rng.integers(-6, 7, size=400)
    vals[vals == 0] = 3
    kw = dict(seed=5, step=3, shift=-2.0, dtau=0.05)
    sty = O.style("integer")
    # no value exceeds the threshold: every spawn is unsafe and every diagonal deposit safe,
    # so only the diagonal step survives; spawns with t = 0 have value 0 and never deposit
    k, v, st = O.step_initiator(model, sty, keys, vals, rule=rule, threshold=1e300, **kw)
    dk, dv, dst = O.step(O.real1d(M, N, 3.0, 0.0), sty, keys, vals, **kw)
    nzk = k[v != 0]
    assert nzk.tolist() == dk.tolist() and v[v != 0].tolist() == dv.tolist()
    assert (v == 0).sum() > 0   # addresses reached only by spawns stay as zero entries
    assert st["deaths"] == dst["deaths"] and st["clones"] == dst["clones"]
    # every value exceeds the threshold: diagonal deposits are initiators, spawns safe,
    # and the integer sums equal the plain step (cancelled sums remain as zero entries)
    k, v, st = O.step_initiator(model, sty, keys, vals, rule=rule, threshold=-1.0, **kw)
    pk, pv, pst = O.step(model, sty, keys, vals, **kw)
    assert k[v != 0].tolist() == pk.tolist() and v[v != 0].tolist() == pv.tolist()
    for name in ("attempts", "spawns", "deaths", "clones", "zombies", "norm"):
        assert st[name] == pst[name]


@pytest.mark.parametrize("rule", RULES)
def test_column_and_list_order_agree_with_initiators(rule):
    model = O.mom1d(8, 6, 2.0, 1.0)
    rng = np.random.default_rng(12)
    keys = np.sort(rng.choice(O.build_basis(8, 6), 300, replace=False))
    vi = rng.integers(-4, 5, size=300)
    vi[vi == 0] = 1
    kw = dict(rule=rule, threshold=2.0, seed=4, step=9, shift=-1.0, dtau=0.02)
    a = O.step_initiator(model, O.style("integer"), keys, vi, order=0, **kw)
    b = O.step_initiator(model, O.style("integer"), keys, vi, order=1, **kw)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2] == b[2]
    vf = rng.normal(size=300) * 3
    a = O.step_initiator(model, O.style("dynamic"), keys, vf, order=0, **kw)
    b = O.step_initiator(model, O.style("dynamic"), keys, vf, order=1, **kw)
    assert np.array_equal(a[0], b[0])
    np.testing.assert_allclose(a[1], b[1], rtol=1e-12)


def test_non_initiator_rule_is_the_plain_step():
    model = O.real1d(8, 8, 3.0, 1.0)
    rng = np.random.default_rng(13)
    keys = np.sort(rng.choice(O.build_basis(8, 8), 300, replace=False))
    vals = rng.exponential(2.0, size=300)
    kw = dict(seed=2, step=1, shift=-1.0, dtau=0.01, order=1)
    a = O.step_initiator(model, O.style("dynamic"), keys, vals, rule="none", threshold=1.0, **kw)
    b = O.step(model, O.style("dynamic"), keys, vals, **kw)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert a[2]["norm"] == pytest.approx(b[2]["norm"], rel=1e-14)  # summed in hash order
    assert {n: x for n, x in a[2].items() if n != "norm"} == {n: x for n, x in b[2].items() if n != "norm"}


def _lomc_shift(model, start, rule, seed, *, steps=6000, dtau=5e-4, target=300.0,
                zeta=0.05):
    """lomc! with DoubleLogUpdate (shiftstrategy.jl:174-181), shift starting at the
    diagonal element of the start address (lomc.jl:188), start vector start => 1."""
    sty = O.style("dynamic")
    keys = np.array([start], dtype=np.uint64)
    vals = np.array([1.0])
    shift = O.diagonal_element(model, start)
    pnorm = 1.0
    xi = zeta ** 2 / 4
    shifts = []
    for step in range(1, steps + 1):
        keys, vals, st = O.step_initiator(model, sty, keys, vals, rule=rule, threshold=1.0,
                                          seed=seed, step=step, shift=shift, dtau=dtau)
        tnorm = st["norm"]
        assert tnorm > 0
        shift = O.double_log_update(shift, tnorm, pnorm, target, zeta=zeta, xi=xi, dtau=dtau)
        pnorm = tnorm
        shifts.append(shift)
    return np.array(shifts)


def test_initiator_energies_below_the_plateau():
    # test/lomc.jl:612-675: BoseFS{10,10} condensate at mode 5, HubbardMom1D u = 4,
    # IsDynamicSemistochastic, DoubleLogUpdate(zeta = 0.05, target 300), dtau 5e-4
    model = O.mom1d(10, 10, 4.0)
    start = O.from_onr((0, 0, 0, 0, 10, 0, 0, 0, 0, 0))
    E0 = -9.251592973178997
    est = {}
    for rule in ("none", "initiator", "simple", "coherent"):
        s = _lomc_shift(model, start, rule, seed=8008)
        br = ST.blocking_analysis(s[1999:])
        est[rule] = (br.mean, br.err)
    E_no, s_no = est["none"]
    E_i1, s_i1 = est["initiator"]
    E_i2, _ = est["simple"]
    E_i3, s_i3 = est["coherent"]
    assert E_no < E0                       # garbage energy without initiators
    assert E_i1 > E0 and E_i2 > E0 and E_i3 > E0
    assert E_i2 > E_i1                     # the simple rule has the largest bias
    assert abs(E_i1 - E_i3) <= max(3 * s_i1, 3 * s_i3)
