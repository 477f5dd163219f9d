"""Rimu's StatsTools error analysis for time series of a run (shift, projected energy).

Host-side analysis of report columns (numpy/scipy); the FCIQMC step itself never calls it.

* ``blocking_analysis``: Flyvbjerg-Petersen blocking with Jonsson's M test,
  ``src/StatsTools/blocking.jl:134-159`` (driver), ``:294-325`` (``blocks_with_m``),
  ``:274-285`` (``mtest``), ``:40-48`` (``BlockingResult`` picks ``mean[1]``,
  ``std_err[k]``), ``src/StatsTools/variances.jl:8-16`` (``autocovariance``).
* ``ratio_of_means``: ``src/StatsTools/ratio_of_means.jl:126-154`` with
  ``ratio_estimators`` (:236-263) and ``x_by_y_linear`` (:167-172): the blocked sample
  means of numerator and denominator, their variances and covariance are sampled as a
  bivariate normal (the reference's MonteCarloMeasurements ``Particles``, 2000 samples by
  default, near-zero variances sampled on their own as in ``particles`` :186-218) and
  divided sample by sample; linear propagation (``sigma_f``) and the denominator's
  coefficient of variation ``delta_y`` are reported beside it.
* ``val_and_errs`` (``src/StatsTools/convenience.jl:61-112``): value and lower/upper error
  bars of n standard errors; for a ratio, the median and the quantiles of the Monte Carlo
  samples at probability ``erf(n / sqrt(2))``.
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


class Particles:
    """Monte Carlo samples of an uncertain number (MonteCarloMeasurements' ``Particles``)."""

    def __init__(self, samples):
        self.samples = np.asarray(samples, dtype=np.float64)

    def pmedian(self) -> float:
        return float(np.median(self.samples))

    def pquantile(self, q):
        with np.errstate(invalid="ignore"):  # samples of a zero-mean denominator hold inf
            return np.quantile(self.samples, q)

    def pmean(self) -> float:
        return float(self.samples.mean())

    def pvar(self) -> float:
        return float(self.samples.var(ddof=1)) if len(self.samples) > 1 else 0.0

    def pextrema(self):
        return float(self.samples.min()), float(self.samples.max())

    def __repr__(self):
        q = self.pquantile([0.16, 0.5, 0.84])
        return f"Particles({q[1]:.6g} +{q[2] - q[1]:.3g} -{q[1] - q[0]:.3g}, n={len(self.samples)})"


@dataclass
class RatioBlockingResult:
    """``RatioBlockingResult`` (ratio_of_means.jl:22-30)."""
    ratio: Particles   # Monte Carlo propagated ratio of the blocked means
    f: float           # ratio of means
    sigma_f: float     # standard error from linear propagation
    delta_y: float     # coefficient of variation of the denominator's mean (< 0.1: normal approx.)
    k: int             # k - 1 blocking steps decorrelate both series
    blocks: int        # data values after blocking
    success: bool      # false if a blocking analysis failed

    def pmedian(self) -> float:
        return self.ratio.pmedian()

    def pquantile(self, q):
        return self.ratio.pquantile(q)

    @property
    def err(self) -> float:
        """Half the Monte Carlo interval of one standard error (val_and_errs with n = 1)."""
        _, lo, hi = val_and_errs(self)
        return 0.5 * (lo + hi)


def _particles_2d(mu, cov, samples, rng):
    """Samples of a bivariate normal; components whose variance is near zero (below
    sqrt(eps) of itself, as ``particles`` in ratio_of_means.jl:186-218) are sampled on their
    own.  Returns an array (2, samples)."""
    mu = np.asarray(mu, dtype=np.float64)
    cov = np.asarray(cov, dtype=np.float64)
    d = np.diag(cov)
    singular = np.array([abs(x) < np.sqrt(np.spacing(abs(x))) if x != 0 else True for x in d])
    out = np.empty((2, samples))
    reg = np.flatnonzero(~singular)
    if len(reg):
        c = cov[np.ix_(reg, reg)]
        w, v = np.linalg.eigh(c)
        a = v * np.sqrt(np.clip(w, 0.0, None))
        out[reg] = mu[reg, None] + a @ rng.standard_normal((len(reg), samples))
    for i in np.flatnonzero(singular):
        out[i] = mu[i] + np.sqrt(max(d[i], 0.0)) * rng.standard_normal(samples)
    return out


def x_by_y_linear(mu_x, mu_y, sigma_x, sigma_y, rho):
    """``x_by_y_linear`` (ratio_of_means.jl:167-172): f = x/y and its standard error by
    linear propagation for correlated normal x, y (covariance rho)."""
    mu_x, mu_y, sigma_x, sigma_y, rho = (np.float64(x) for x in (mu_x, mu_y, sigma_x, sigma_y, rho))
    with np.errstate(divide="ignore", invalid="ignore"):  # a zero mean gives inf/nan, as in Julia
        f = mu_x / mu_y
        v = (sigma_x / mu_y) ** 2 + (mu_x * sigma_y / mu_y ** 2) ** 2 - 2 * rho * mu_x / mu_y ** 3
    return float(f), float(np.sqrt(v)) if v >= 0 else float("nan")


def ratio_estimators(x, y, k=None, corrected=True, mc_samples=2000, rng=None):
    """``ratio_estimators`` (ratio_of_means.jl:236-271): k - 1 blocking steps, then the
    Monte Carlo ratio of the sample means, its linear propagation and ``delta_y``.
    Returns (Particles, f, sigma_f, delta_y, n)."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    if k is not None:
        for _ in range(k - 1):
            x, y = _blocker(x), _blocker(y)
    n = len(x)
    c = 1 if corrected else 0
    mx, my = float(x.mean()), float(y.mean())
    var_x = float(np.dot(x - mx, x - mx) / (n - c)) / n
    var_y = float(np.dot(y - my, y - my) / (n - c)) / n
    rho = float(np.dot(x - mx, y - my) / (n - c)) / n
    rng = np.random.default_rng(0) if rng is None else rng
    xy = _particles_2d([mx, my], [[var_x, rho], [rho, var_y]], int(mc_samples), rng)
    with np.errstate(divide="ignore", invalid="ignore"):
        r = Particles(xy[0] / xy[1])
    f, sigma_f = x_by_y_linear(mx, my, np.sqrt(var_x), np.sqrt(var_y), rho)
    delta_y = float(np.sqrt(var_y) / my) if my != 0 else float("inf")
    return r, f, sigma_f, delta_y, n


def ratio_of_means(num, denom, alpha=0.01, corrected=True, skip=0, mc_samples=2000,
                   rng=None) -> RatioBlockingResult:
    """``ratio_of_means(num, denom; alpha, corrected, mc_samples, skip)``
    (ratio_of_means.jl:126-154): blocking analyses of both series pick k (the larger of
    the two M-test results; ``success`` is false when either fails), then the blocked means
    are propagated through the ratio by Monte Carlo.  ``f`` is the ratio of the unblocked
    means.  |delta_y| >= 0.1 means the linear error (``sigma_f``) is not to be trusted."""
    num = np.asarray(num, dtype=np.float64)[skip:]
    denom = np.asarray(denom, dtype=np.float64)[skip:]
    kn = mtest(blocks_with_m(num, corrected)["mj"], alpha)
    kd = mtest(blocks_with_m(denom, corrected)["mj"], alpha)
    success = kn >= 0 and kd >= 0
    k = max(kn, kd)
    r, _, sigma_f, delta_y, blocks = ratio_estimators(num, denom, k, corrected, mc_samples, rng)
    with np.errstate(divide="ignore", invalid="ignore"):
        f = float(np.float64(num.mean()) / np.float64(denom.mean()))
    return RatioBlockingResult(r, f, sigma_f, delta_y, k, blocks, success)


def val_and_errs(r, n=1, p=None):
    """``val_and_errs`` (convenience.jl:61-112): (value, lower error, upper error) of n
    standard errors, or of the central probability p.  A BlockingResult is normal (mean,
    n * err); a RatioBlockingResult gives its Monte Carlo median and quantiles."""
    from math import erf, sqrt
    from scipy.stats import norm
    if p is None:
        p = erf(n / sqrt(2))
    c1, c2 = (1 - p) / 2, (1 + p) / 2
    if isinstance(r, RatioBlockingResult):
        q1, q2, q3 = r.ratio.pquantile([c1, 0.5, c2])
        return float(q2), float(q2 - q1), float(q3 - q2)
    if isinstance(r, BlockingResult):
        lo = norm.ppf(c1, loc=r.mean, scale=r.err) if r.err > 0 else r.mean
        hi = norm.ppf(c2, loc=r.mean, scale=r.err) if r.err > 0 else r.mean
        return r.mean, float(r.mean - lo), float(hi - r.mean)
    return float(r), 0.0, 0.0


def within(r, x, n=2) -> bool:
    """Whether x lies inside the n-standard-error interval of r (val_and_errs)."""
    v, lo, hi = val_and_errs(r, n=n)
    return v - lo <= x <= v + hi


def shift_estimator(df, skip=0, **kw) -> BlockingResult:
    return blocking_analysis(np.asarray(df["shift"]), skip=skip, **kw)


def projected_energy(df, skip=0, hproj="hproj", vproj="vproj", **kw) -> RatioBlockingResult:
    return ratio_of_means(np.asarray(df[hproj]), np.asarray(df[vproj]), skip=skip, **kw)


def rayleigh_replica_estimator(df, op_name="Op1", vec_name="dot", skip=0, Anorm=1.0,
                               **kw) -> RatioBlockingResult:
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
