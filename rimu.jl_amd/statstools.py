"""Rimu's StatsTools error analysis for time series of a run (shift, projected energy).

Host-side analysis of report columns (numpy/scipy); the FCIQMC step itself never calls it.

* ``blocking_analysis``: Flyvbjerg-Petersen blocking with Jonsson's M test,
  ``src/StatsTools/blocking.jl:134-159`` (driver), ``:294-325`` (``blocks_with_m``),
  ``:274-285`` (``mtest``), ``:40-48`` (``BlockingResult`` picks ``mean[1]``,
  ``std_err[k]``), ``src/StatsTools/variances.jl:8-16`` (``autocovariance``).
* ``ratio_of_means``: ``src/StatsTools/ratio_of_means.jl:126-154``.  The reference
  propagates the blocked means, variances and covariance through the ratio with
  MonteCarloMeasurements particles; this restatement uses the same blocked quantities with
  linear (delta-method) propagation of the error.
* ``shift_estimator`` / ``projected_energy``: ``src/StatsTools/reweighting.jl:721-743``.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
from scipy.stats import chi2


@dataclass
class BlockingResult:
    mean: float
    err: float
    err_err: float
    p_cov: float
    k: int
    blocks: int


def smoothen(noisy, b):
    """Moving average over b elements with periodic wrap (StatsTools/growth_witness.jl:66-77)."""
    noisy = np.asarray(noisy, dtype=np.float64)
    n = len(noisy)
    offset = b // 2 + 1
    out = np.zeros(n)
    for j in range(1, b + 1):
        out += np.roll(noisy, -(j - offset)) / b
    return out


def _blocker(v):
    n = len(v) // 2
    return (v[0:2 * n:2] + v[1:2 * n:2]) / 2


def blocks_with_m(v, corrected=True):
    v = np.asarray(v, dtype=np.float64)
    n_steps = int(np.floor(np.log2(len(v))))
    out = {"blocks": [], "mean": [], "std_err": [], "std_err_err": [], "p_cov": [], "mj": []}
    c = 1 if corrected else 0
    for _ in range(n_steps):
        n = len(v)
        mean_v = v.mean()
        d = v - mean_v
        variance = float(np.dot(d, d) / (n - c))
        gamma = float(np.dot(d[:-1], d[1:]) / (n - c))
        # a constant series gives 0/0 = NaN, which never passes the M test (as in Julia)
        mj = n * ((n - 1) * variance / n ** 2 + gamma) ** 2 / variance ** 2 if variance > 0 else np.nan
        se = np.sqrt(variance / n)
        out["blocks"].append(n)
        out["mean"].append(mean_v)
        out["std_err"].append(se)
        out["std_err_err"].append(se / np.sqrt(2 * (n - 1)))
        out["p_cov"].append(float(np.dot(d, d) / (n - c)) / n)
        out["mj"].append(mj)
        v = _blocker(v)
    return out


def mtest(mj, alpha=0.01):
    m = np.cumsum(np.asarray(mj)[::-1])[::-1]
    k = 1
    while k <= len(m) - 1:
        if m[k - 1] < chi2.isf(alpha, k):  # cquantile(Chisq(k), alpha)
            return k
        k += 1
    return -1


def blocking_analysis(v, alpha=0.01, corrected=True, skip=0) -> BlockingResult:
    v = np.asarray(v, dtype=np.float64)[skip:]
    if len(v) < 2:
        return BlockingResult(float(v[0]) if len(v) else 0.0, np.nan, np.nan, np.nan, -1, len(v))
    nt = blocks_with_m(v, corrected)
    k = mtest(nt["mj"], alpha)
    if k < 0:
        return BlockingResult(nt["mean"][0], np.nan, np.nan, np.nan, k, 0)
    return BlockingResult(nt["mean"][0], nt["std_err"][k - 1], nt["std_err_err"][k - 1],
                          nt["p_cov"][k - 1], k, nt["blocks"][k - 1])


@dataclass
class RatioResult:
    ratio: float
    err: float
    k: int
    blocks: int


def ratio_of_means(num, denom, alpha=0.01, corrected=True, skip=0) -> RatioResult:
    num = np.asarray(num, dtype=np.float64)[skip:]
    denom = np.asarray(denom, dtype=np.float64)[skip:]
    bn = blocks_with_m(num, corrected)
    bd = blocks_with_m(denom, corrected)
    k = max(mtest(bn["mj"], alpha), mtest(bd["mj"], alpha))
    if k < 0 or min(mtest(bn["mj"], alpha), mtest(bd["mj"], alpha)) < 0:
        return RatioResult(float(num.mean() / denom.mean()), np.nan, -1, 0)
    # blocked series at level k
    x, y = num.copy(), denom.copy()
    for _ in range(k - 1):
        x, y = _blocker(x), _blocker(y)
    n = len(x)
    mx, my = x.mean(), y.mean()
    vx = np.var(x, ddof=1) / n
    vy = np.var(y, ddof=1) / n
    cxy = np.dot(x - mx, y - my) / (n - 1) / n
    r = mx / my
    err = abs(r) * np.sqrt(max(vx / mx ** 2 + vy / my ** 2 - 2 * cxy / (mx * my), 0.0))
    return RatioResult(float(num.mean() / denom.mean()), float(err), k, n)


def shift_estimator(df, skip=0, **kw) -> BlockingResult:
    return blocking_analysis(np.asarray(df["shift"]), skip=skip, **kw)


def projected_energy(df, skip=0, hproj="hproj", vproj="vproj", **kw) -> RatioResult:
    return ratio_of_means(np.asarray(df[hproj]), np.asarray(df[vproj]), skip=skip, **kw)


def rayleigh_replica_estimator(df, op_name="Op1", vec_name="dot", skip=0, Anorm=1.0,
                               **kw) -> RatioResult:
    """``rayleigh_replica_estimator(df; op_name, vec_name, skip, Anorm)`` without reweighting
    (h = 0; reweighting.jl:526-583): the ratio of means of the operator overlaps summed over
    replica pairs, ``sum_{a<b} c{a}_{op}_c{b}``, to the vector overlaps ``sum c{a}_{vec}_c{b}``."""
    n = sum(1 for c in df.columns if c.startswith("norm_"))
    num = 0.0
    den = 0.0
    for a in range(1, n + 1):
        for b in range(a + 1, n + 1):
            num = num + np.asarray(df[f"c{a}_{op_name}_c{b}"], dtype=float) * Anorm
            den = den + np.asarray(df[f"c{a}_{vec_name}_c{b}"], dtype=float)
    return ratio_of_means(np.asarray(num)[skip:], np.asarray(den)[skip:], **kw)
