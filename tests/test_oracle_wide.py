"""CPU tests that pin the oracle on two-chunk addresses (64 < B = N + M - 1 <= 128) before
the device's 128-bit path is checked against it.

Anchors independent of the oracle's bit arithmetic:
* the layout fixed by bitstring.jl:464-508 (mode 1 in the lowest bits of the last chunk,
  chunks continuing upward): near_uniform(BoseFS{50,50}) is the 99-bit pattern 0b0101...01,
  and every address round-trips through its occupation numbers;
* Hamiltonian columns built from the second-quantised operators on occupation tuples
  (a+ a hops for HubbardReal1D, a+ a+ a a momentum transfers for HubbardMom1D) equal the
  oracle's diagonal and off-diagonal elements;
* a deterministic step of a two-chunk vector equals T v built from those columns, and the
  list order of deposits gives the column order's result (integer walkers, exactly).
"""
import math

import numpy as np
import pytest

import oracle as O


def _random_onr(M, N, rng):
    B = N + M - 1
    cuts = np.sort(rng.choice(B, M - 1, replace=False))
    return tuple(int(x) for x in np.diff(np.concatenate([[-1], cuts, [B]])) - 1)


def _real_column(onr, u, t):
    M = len(onr)
    diag = u / 2 * sum(n * (n - 1) for n in onr)
    off = {}
    for i in range(M):
        if onr[i] == 0:
            continue
        for j in ((i + 1) % M, (i - 1) % M):
            s = list(onr)
            amp = math.sqrt(s[i])
            s[i] -= 1
            s[j] += 1
            amp *= math.sqrt(s[j])
            off[tuple(s)] = off.get(tuple(s), 0.0) - t * amp
    return diag, off


def _mom_column(onr, u, kes):
    M = len(onr)
    diag = sum(kes[i] * n for i, n in enumerate(onr))
    off = {}
    for p in range(M):
        if onr[p] == 0:
            continue
        for q in range(M):
            s1 = list(onr)
            f = math.sqrt(s1[p])
            s1[p] -= 1
            if s1[q] == 0:
                continue
            f *= math.sqrt(s1[q])
            s1[q] -= 1
            for k in range(M):
                s = list(s1)
                s[(q - k) % M] += 1
                g = f * math.sqrt(s[(q - k) % M])
                s[(p + k) % M] += 1
                g *= math.sqrt(s[(p + k) % M])
                amp = u / (2 * M) * g
                key = tuple(s)
                if key == tuple(onr):
                    diag += amp
                else:
                    off[key] = off.get(key, 0.0) + amp
    return diag, off


def _oracle_column(model, onr):
    bits = O.from_onr(onr)
    d = O.diagonal_element(model, bits)
    off = {}
    for c in range(1, O.num_offdiagonals(model, bits) + 1):
        nb, v = O.get_offdiagonal(model, bits, c)
        t = O.to_onr(nb, model.M)
        off[t] = off.get(t, 0.0) + v
    return d, off


def test_layout_known_answer_and_round_trip():
    bits = O.from_onr((1,) * 50)                  # near_uniform(BoseFS{50,50})
    assert bits == sum(1 << (2 * i) for i in range(50)) and bits.bit_length() == 99
    assert O.from_onr((64,) + (0,) * 63) == (1 << 64) - 1   # 64 bosons in mode 1, B = 127
    assert O.from_onr((0,) * 63 + (64,)) == ((1 << 64) - 1) << 63
    rng = np.random.default_rng(1)
    for M, N in ((50, 50), (40, 30), (64, 64), (100, 28), (65, 64)):
        for _ in range(50):
            onr = _random_onr(M, N, rng)
            b = O.from_onr(onr)
            assert b < (1 << (N + M - 1)) and bin(b).count("1") == N
            assert O.to_onr(b, M) == onr


@pytest.mark.parametrize("M,N", [(50, 50), (100, 28), (64, 64)])
def test_real1d_columns_match_second_quantisation(M, N):
    rng = np.random.default_rng(M + N)
    model = O.real1d(M, N, 4.0, 1.3)
    for _ in range(20):
        onr = _random_onr(M, N, rng)
        d, off = _oracle_column(model, onr)
        bd, boff = _real_column(onr, 4.0, 1.3)
        assert d == pytest.approx(bd, rel=1e-14)
        assert off.keys() == boff.keys()
        for k in off:
            assert off[k] == pytest.approx(boff[k], rel=1e-13)


@pytest.mark.parametrize("M,N", [(40, 30), (30, 50)])
def test_mom1d_columns_match_second_quantisation(M, N):
    rng = np.random.default_rng(M * N)
    model = O.mom1d(M, N, 3.0, 1.0)
    _, kes = O.mom1d_grid(M, 1.0)
    for _ in range(3):
        onr = _random_onr(M, N, rng)
        d, off = _oracle_column(model, onr)
        bd, boff = _mom_column(onr, 3.0, kes)
        assert d == pytest.approx(bd, rel=1e-12)
        boff = {k: v for k, v in boff.items() if abs(v) > 1e-13}
        off = {k: v for k, v in off.items() if abs(v) > 1e-13}
        assert off.keys() == boff.keys()
        for k in off:
            assert off[k] == pytest.approx(boff[k], rel=1e-12)


def test_deterministic_step_equals_brute_force_columns():
    M = N = 50
    u, t, shift, dtau = 4.0, 1.0, -2.0, 0.05
    model = O.real1d(M, N, u, t)
    rng = np.random.default_rng(5)
    onrs = sorted({_random_onr(M, N, rng) for _ in range(60)}, key=O.from_onr)
    keys = O.pack_keys([O.from_onr(o) for o in onrs], 2)
    vals = rng.normal(size=len(onrs))
    ok, ov, st = O.step(model, O.style("deterministic"), keys, vals, seed=1, step=1, shift=shift,
                        dtau=dtau, order=0)
    want = {}
    for onr, v in zip(onrs, vals):
        d, off = _real_column(onr, u, t)
        want[onr] = want.get(onr, 0.0) + (1 - dtau * (d - shift)) * v
        for k, a in off.items():
            want[k] = want.get(k, 0.0) - dtau * a * v
    got = {O.to_onr(b, M): v for b, v in zip(O.unpack_keys(ok), ov.tolist())}
    assert got.keys() == {k for k, v in want.items() if v != 0}
    for k, v in got.items():
        assert v == pytest.approx(want[k], rel=1e-12, abs=1e-14)
    assert ok.shape == (len(got), 2)


@pytest.mark.parametrize("kind", ["real", "mom"])
def test_integer_column_and_list_order_agree(kind):
    M, N = (50, 50) if kind == "real" else (40, 30)
    model = O.real1d(M, N, 6.0, 1.0) if kind == "real" else O.mom1d(M, N, 4.0, 1.0)
    rng = np.random.default_rng(9)
    ints = sorted({O.from_onr(_random_onr(M, N, rng)) for _ in range(400)})
    keys = O.pack_keys(ints, 2)
    vals = rng.integers(-9, 10, size=len(ints))
    vals[vals == 0] = 3
    kw = dict(seed=3, step=2, shift=-1.0, dtau=0.01)
    a = O.step(model, O.style("integer"), keys, vals, order=0, **kw)
    b = O.step(model, O.style("integer"), keys, vals, order=1, **kw)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert a[2]["attempts"] == b[2]["attempts"] > 0 and len(a[0]) > len(ints)
