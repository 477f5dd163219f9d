"""CPU tests that pin the oracle's HubbardRealSpace (BoseFS, FermiFS, two-component
CompositeFS on CubicGrid geometries) before the device is checked against it.

Anchors:
* the reference's own tests, test/Hamiltonians.jl:263-460: channel counts of
  near_uniform(FermiFS{3,12}) on four geometries, exact energies of 1D and 2D fermions
  (-14, -9, -8, -4, -6, -8, -10, -16), equalities of 1D boson, two-component and trap
  energies, and HubbardRealSpace = HubbardReal1D in 1D;
* columns built here from second-quantised hopping on occupation tuples (Jordan-Wigner
  signs for fermions, sqrt factors for bosons) over a lattice bond list written out
  independently of the oracle's neighbour table;
* a deterministic step equal to T v from those columns (zero-valued channels at hard walls
  and Pauli-blocked hops deposit nothing), and integer steps in column and list order equal.
"""
import itertools
import math

import numpy as np
import pytest

import oracle as O


def exact_energy(model, start):
    basis = O.bfs_basis(model, start)
    return float(O.ed_energies(model, basis, 1)[0])


def krylov_energy(model, start):
    """The energy eigsolve(H, DVec(start => 1.0), 1, :SR) converges to (the reference's
    exact_energy, test/Hamiltonians.jl:13-17): the lowest eigenvalue whose eigenspace has
    weight on the start vector."""
    basis = O.bfs_basis(model, start)
    H = O.sparse_hamiltonian(model, basis).toarray()
    w, V = np.linalg.eigh(0.5 * (H + H.T))
    j = int(np.searchsorted(basis, np.uint64(start)))
    weight = V[j, :] ** 2
    i = 0
    while i < len(w):
        k = i
        while k + 1 < len(w) and w[k + 1] - w[i] < 1e-9:
            k += 1
        if weight[i:k + 1].sum() > 1e-12:
            return float(w[i])
        i = k + 1
    raise AssertionError("start vector has no weight")


def fermi(n, M):
    return tuple([1] * n + [0] * (M - n))  # near_uniform(FermiFS{n,M})


# ---------------------------------------------------------------------------------------
# independent brute force
# ---------------------------------------------------------------------------------------
def bonds(dims, fold):
    """Directed moves (src, dst, direction index) of a CubicGrid, from coordinates."""
    D = len(dims)
    coords = list(itertools.product(*[range(n) for n in reversed(dims)]))
    coords = [tuple(reversed(c)) for c in coords]            # first dimension fastest
    index = {c: i for i, c in enumerate(coords)}
    out = {}
    for c in coords:
        for d in range(2 * D):
            y = list(c)
            y[d % D] += 1 if d < D else -1
            if fold[d % D]:
                y[d % D] %= dims[d % D]
            elif not 0 <= y[d % D] < dims[d % D]:
                out[(index[c], d)] = None
                continue
            out[(index[c], d)] = index[tuple(y)]
    return out


def brute_column(comps, dims, fold, t, u, pot, onrs):
    """Diagonal and {onrs': value} of H's column at onrs (lists of occupation tuples)."""
    M = math.prod(dims)
    nb = bonds(dims, fold)
    D = len(dims)
    diag = 0.0
    C = len(comps)
    for c in range(C):
        if comps[c][0] == "bose":
            diag += u[c][c] / 2 * sum(n * (n - 1) for n in onrs[c])
    if C == 2:
        diag += u[0][1] * sum(a * b for a, b in zip(onrs[0], onrs[1]))
    for c in range(C):
        diag += sum(n * pot[c][i] for i, n in enumerate(onrs[c]))
    off = {}
    for c in range(C):
        onr = onrs[c]
        for src in range(M):
            if onr[src] == 0:
                continue
            for d in range(2 * D):
                dst = nb[(src, d)]
                if dst is None:
                    continue
                new = list(onr)
                if comps[c][0] == "fermi":
                    if new[dst] and dst != src:
                        continue
                    sign = (-1) ** sum(new[:src])
                    new[src] = 0
                    sign *= (-1) ** sum(new[:dst])
                    new[dst] = 1
                    amp = -t[c] * sign
                else:
                    amp = math.sqrt(new[src])
                    new[src] -= 1
                    new[dst] += 1
                    amp *= math.sqrt(new[dst])
                    amp = -t[c] * amp
                key = tuple(tuple(x) for x in (onrs[:c] + [tuple(new)] + onrs[c + 1:]))
                off[key] = off.get(key, 0.0) + amp
    return diag, off


def oracle_column(model, key):
    d = O.diagonal_element(model, key)
    off = {}
    for ch in range(1, O.num_offdiagonals(model, key) + 1):
        k, v = O.get_offdiagonal(model, key, ch)
        if v == 0:
            assert k == key   # a hard wall or a Pauli-blocked hop returns the address itself
            continue
        o = O.rs_onrs(model, k)
        off[o] = off.get(o, 0.0) + v
    return d, off


def random_onr(kind, n, M, rng):
    if kind == "fermi":
        x = [0] * M
        for i in rng.choice(M, n, replace=False):
            x[i] = 1
        return tuple(x)
    cuts = np.sort(rng.choice(n + M - 1, M - 1, replace=False))
    return tuple(int(x) for x in np.diff(np.concatenate([[-1], cuts, [n + M - 1]])) - 1)


# ---------------------------------------------------------------------------------------
# the reference's known answers
# ---------------------------------------------------------------------------------------
@pytest.mark.parametrize("dims,fold,want", [((3, 4), (True, True), 6), ((4, 3), (True, True), 8),
                                            ((3, 4), (False, False), 3), ((4, 3), (False, False), 4)])
def test_offdiagonal_counts(dims, fold, want):
    m = O.real_space(dims, [("fermi", 3)], fold=fold)
    od = O.offdiagonals(m, O.rs_key(m, [fermi(3, 12)]))
    assert len(od) == 12 and sum(1 for _, v in od if v != 0) == want


def test_1d_bosons_equal_hubbard_real_1d():
    onr = (1, 1, 1, 1, 1, 0)
    rs = O.real_space((6,), [("bose", 5)], t=[3], u=[[2]])
    r1 = O.real1d(6, 5, 2.0, 3.0)
    key = O.rs_key(rs, [onr])
    assert key == O.from_onr(onr)
    assert exact_energy(rs, key) == pytest.approx(exact_energy(r1, key), rel=1e-12)
    for b in O.build_basis(6, 5).tolist():   # the same columns, element for element
        assert O.diagonal_element(rs, b) == O.diagonal_element(r1, b)
        a = sorted((k, v) for k, v in O.offdiagonals(rs, b))
        c = sorted((k, v) for k, v in O.offdiagonals(r1, b))
        assert a == c


def test_1d_two_component_bosons():
    b3, b1 = (1, 1, 1, 0, 0, 0), (1, 0, 0, 0, 0, 0)
    cases = [([("bose", 3), ("bose", 1)], [b3, b1], [1, 4], [[2, 3], [3, 0]]),
             ([("bose", 3), ("fermi", 1)], [b3, b1], [1, 4], [[2, 3], [3, 0]]),
             ([("bose", 1), ("bose", 3)], [b1, b3], [4, 1], [[0, 3], [3, 2]]),
             ([("fermi", 1), ("bose", 3)], [b1, b3], [4, 1], [[0, 3], [3, 2]])]
    E = []
    for comps, onrs, t, u in cases:
        m = O.real_space((6,), comps, t=t, u=u)
        E.append(exact_energy(m, O.rs_key(m, onrs)))
    assert E[0] == pytest.approx(E[1], rel=1e-10)
    assert E[1] == pytest.approx(E[2], rel=1e-10)
    assert E[2] == pytest.approx(E[3], rel=1e-10)


def test_1d_fermions():
    m = O.real_space((6,), [("fermi", 3)], t=[3.5])
    assert exact_energy(m, O.rs_key(m, [fermi(3, 6)])) == pytest.approx(-14, rel=1e-10)
    comps = [("fermi", 4), ("fermi", 2)]
    onrs = [fermi(4, 6), fermi(2, 6)]
    E = {}
    for name, u in [("free", [[0, 0], [0, 0]]), ("rep", [[0, 1], [1, 0]]), ("att", [[0, -1], [-1, 0]])]:
        m = O.real_space((6,), comps, t=[1, 2], u=u)
        E[name] = exact_energy(m, O.rs_key(m, onrs))
    assert E["free"] == pytest.approx(-9, rel=1e-10)
    assert E["rep"] > -9 and E["att"] < -9


def test_1d_trap():
    onr = (1, 2, 3, 4)
    m = O.real_space((4,), [("bose", 10)], t=[3], u=[[2]], v=[[4]])
    key = O.rs_key(m, [onr])
    # HubbardReal1DEP: HubbardReal1D plus v x^2 n with x = i - 1 - fld(M, 2)
    basis = O.bfs_basis(m, key)
    H1 = O.sparse_hamiltonian(O.real1d(4, 10, 2.0, 3.0), basis).toarray()
    x = np.arange(4) - 2
    for j, b in enumerate(basis.tolist()):
        H1[j, j] += sum(4.0 * xi * xi * n for xi, n in zip(x, O.to_onr(b, 4)))
    want = np.linalg.eigvalsh(H1)[0]
    assert exact_energy(m, key) == pytest.approx(want, rel=1e-12)
    b3, b1 = (1, 1, 1, 0, 0, 0), (1, 0, 0, 0, 0, 0)
    m3 = O.real_space((6,), [("bose", 3), ("bose", 1)], v=[[1], [4]], u=[[2, 3], [3, 0]])
    m4 = O.real_space((6,), [("bose", 1), ("bose", 3)], v=[[4], [1]], u=[[0, 3], [3, 2]])
    assert exact_energy(m3, O.rs_key(m3, [b3, b1])) == pytest.approx(
        exact_energy(m4, O.rs_key(m4, [b1, b3])), rel=1e-10)


@pytest.mark.parametrize("n,want", [(1, -8), (2, -8), (3, -8), (4, 0)])
def test_2d_fermions_2x2(n, want):
    m = O.real_space((2, 2), [("fermi", n)], t=[2])
    assert exact_energy(m, O.rs_key(m, [fermi(n, 4)])) == pytest.approx(want, abs=1e-10)


@pytest.mark.parametrize("onr,want", [(fermi(1, 16), -4), (fermi(2, 16), -6), (fermi(3, 16), -8),
                                      ((1, 0, 1, 0, 1, 0, 1, 0, 1, 0, 0, 0, 0, 0, 0, 0), -10)])
def test_2d_fermions_4x4(onr, want):
    """The reference's values are eigsolve's from the start address.  For five fermions
    that is -10, the lowest level in the start vector's symmetry sector: the closed-shell
    ground state at -4 + 4 * (-2) = -12 has no weight on it."""
    m = O.real_space((4, 4), [("fermi", sum(onr))])
    key = O.rs_key(m, [onr])
    assert krylov_energy(m, key) == pytest.approx(want, rel=1e-10)
    free = sorted(-2 * (math.cos(2 * math.pi * a / 4) + math.cos(2 * math.pi * b / 4))
                  for a in range(4) for b in range(4))
    assert exact_energy(m, key) == pytest.approx(sum(free[:sum(onr)]), abs=1e-10)


def test_2d_two_component_fermions():
    onrs = [fermi(3, 9), fermi(2, 9)]
    E = {}
    for name, u in [("free", [[0, 0], [0, 0]]), ("rep", [[0, 1], [1, 0]]), ("att", [[0, -1], [-1, 0]])]:
        m = O.real_space((3, 3), [("fermi", 3), ("fermi", 2)], t=[1, 2], u=u)
        E[name] = exact_energy(m, O.rs_key(m, onrs))
    assert E["free"] == pytest.approx(-16, rel=1e-10)
    assert E["rep"] > -16 and E["att"] < -16


# ---------------------------------------------------------------------------------------
# second-quantised brute force
# ---------------------------------------------------------------------------------------
GEOMS = [((4, 4), (True, True)), ((3, 5), (False, False)), ((4, 3), (False, True)),
         ((3, 2, 2), (True, False, True)), ((7,), (False,)), ((8, 8), (True, True))]
COMPS = [[("fermi", 3), ("fermi", 4)], [("bose", 4)], [("fermi", 5)], [("bose", 3), ("fermi", 2)],
         [("fermi", 2), ("bose", 3)], [("bose", 2), ("bose", 3)]]


@pytest.mark.parametrize("gi", range(len(GEOMS)))
@pytest.mark.parametrize("ci", range(len(COMPS)))
def test_columns_match_second_quantisation(gi, ci):
    dims, fold = GEOMS[gi]
    comps = COMPS[ci]
    M = math.prod(dims)
    if sum(M if k == "fermi" else n + M - 1 for k, n in comps) > 128:
        pytest.skip("more than 128 bits")
    rng = np.random.default_rng(100 * gi + ci)
    C = len(comps)
    t = [1.3, -0.7][:C]
    u = [[2.0, 1.5], [1.5, 0.5]] if C == 2 else [[2.0]]
    v = rng.normal(size=(C, len(dims))).tolist()
    m = O.real_space(dims, comps, t=t, u=[r[:C] for r in u[:C]], v=v, fold=fold)
    pot = [O.trap_potential(dims, v[c]) for c in range(C)]
    for _ in range(15):
        onrs = [random_onr(k, n, M, rng) for k, n in comps]
        key = O.rs_key(m, onrs)
        assert O.rs_onrs(m, key) == tuple(onrs)
        d, off = oracle_column(m, key)
        bd, boff = brute_column(comps, dims, fold, t, u, pot, onrs)
        assert d == pytest.approx(bd, rel=1e-12, abs=1e-12)
        boff = {k: x for k, x in boff.items() if x != 0}
        off = {k: x for k, x in off.items() if x != 0}
        assert off.keys() == boff.keys()
        for k in off:
            assert off[k] == pytest.approx(boff[k], rel=1e-14)


def test_deterministic_step_equals_brute_force():
    dims, fold = (4, 3), (False, True)
    comps = [("fermi", 4), ("fermi", 3)]
    t, u = [1.0, 0.8], [[0.0, 4.0], [4.0, 0.0]]
    m = O.real_space(dims, comps, t=t, u=u, fold=fold)
    pot = [[0.0] * 12, [0.0] * 12]
    rng = np.random.default_rng(2)
    onrs_list = sorted({tuple(random_onr(k, n, 12, rng) for k, n in comps) for _ in range(40)},
                       key=lambda o: O.rs_key(m, o))
    keys = np.array([O.rs_key(m, o) for o in onrs_list], dtype=np.uint64)
    vals = rng.normal(size=len(keys))
    shift, dtau = -1.0, 0.02
    ok, ov, st = O.step(m, O.style("deterministic"), keys, vals, seed=1, step=1, shift=shift,
                        dtau=dtau, order=0)
    want = {}
    for o, x in zip(onrs_list, vals):
        d, off = brute_column(comps, dims, fold, t, u, pot, list(o))
        want[o] = want.get(o, 0.0) + (1 - dtau * (d - shift)) * x
        for k, a in off.items():
            want[k] = want.get(k, 0.0) - dtau * a * x
    got = {O.rs_onrs(m, int(k)): x for k, x in zip(ok.tolist(), ov.tolist())}
    assert got.keys() == {k for k, x in want.items() if x != 0}
    for k, x in got.items():
        assert x == pytest.approx(want[k], rel=1e-12, abs=1e-14)
    assert st["attempts"] == sum(O.num_offdiagonals(m, int(k)) for k in keys)


@pytest.mark.parametrize("ci", [0, 3])
def test_integer_column_and_list_order_agree(ci):
    comps = COMPS[ci]
    dims, fold = (4, 4), (True, False)
    m = O.real_space(dims, comps, u=[[1.0, 2.0], [2.0, 0.0]][:len(comps)], fold=fold)
    rng = np.random.default_rng(ci)
    keys = np.array(sorted({O.rs_key(m, [random_onr(k, n, 16, rng) for k, n in comps])
                            for _ in range(500)}), dtype=np.uint64)
    vals = rng.integers(-9, 10, size=len(keys))
    vals[vals == 0] = 2
    kw = dict(seed=5, step=3, shift=-2.0, dtau=0.02)
    a = O.step(m, O.style("integer"), keys, vals, order=0, **kw)
    b = O.step(m, O.style("integer"), keys, vals, order=1, **kw)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert a[2]["attempts"] > 0 and len(a[0]) > len(keys) // 2
