"""The TMCMC sampler (gh_tmcmc, host code) pinned by what fixes it
independently of its implementation:

1. the annealing exponent: the coefficient of variation of the plausibility
   weights at p_{j+1}, computed here with numpy, equals TolCOV; flat
   likelihoods give the reference's fallback step p_j + 1e-5; the exponent is
   capped at 1;
2. a Gaussian likelihood under a uniform prior: the final population's mean
   and spread match the analytic posterior, and the log evidence matches the
   closed form log(2 pi sigma^2 / area);
3. the glioma parameter map (fitfun_glioma.c:118-152) and the Likelihood.txt
   round trip, restated with numpy and Python's own text formatting;
4. failure handling: non-finite log-likelihoods become the -1e12 penalty,
   wrong counts and calls after the end are errors.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import gliomasolver_amd as gh
from gliomasolver_amd._capi import GHError


def cov_of(f, x, p):
    w = np.exp((f - f.max()) * (x - p))
    q = w / w.sum()
    return q.std(ddof=1) / q.mean()


@pytest.mark.parametrize("p", [0.0, 0.013, 0.2])
def test_next_exponent_hits_target_cov(p):
    rng = np.random.default_rng(1)
    f = rng.normal(-50, 20, 3000)
    pn, cov = gh.tmcmc_next_exponent(f, p, 1.0)
    assert pn > p
    assert abs(cov_of(f, pn, p) - 1.0) < 1e-5
    assert abs(cov - 1.0) < 1e-5


def test_next_exponent_cap_and_flat_fallback():
    f = np.array([-1.0, -2.0, -1.5, -400.0] * 10)
    pn, _ = gh.tmcmc_next_exponent(f * 1e-6, 0.5, 1.0)       # nearly flat: CoV never reaches 1 below p = 4
    assert pn == pytest.approx(0.5 + 1e-5, abs=1e-12) or pn == 1.0
    pn, _ = gh.tmcmc_next_exponent(np.zeros(50), 0.3, 1.0)     # flat: the reference's tiny step
    assert pn == pytest.approx(0.3 + 1e-5, abs=1e-15)
    rng = np.random.default_rng(2)
    pn, _ = gh.tmcmc_next_exponent(rng.normal(0, 1e-3, 500), 0.99, 1.0)   # root beyond 1: capped
    assert pn == 1.0


def run_gaussian(pop, seed, mu, sigma, lo=-5.0, hi=5.0):
    t = gh.Tmcmc([lo, lo], [hi, hi], pop, max_stages=60, min_chain=0, max_chain=1, tol_cov=1.0, bbeta=0.2,
                 seed=seed)
    gens = 0
    while True:
        x = t.candidates()
        if len(x) == 0:
            break
        f = -0.5 * (((x - mu) / sigma) ** 2).sum(axis=1)
        t.submit(f)
        gens += 1
        assert gens < 200
    st = t.state()
    pop_x, pop_f = t.population()
    t.close()
    return st, pop_x, pop_f


def test_gaussian_posterior_and_evidence():
    mu = np.array([1.0, -0.5])
    sigma = 0.5
    st, x, f = run_gaussian(3000, 7, mu, sigma)
    assert st["done"] and st["p"] == 1.0
    assert np.all(np.abs(x.mean(axis=0) - mu) < 0.05)
    assert np.all(np.abs(x.std(axis=0) / sigma - 1) < 0.15)
    exact = math.log(2 * math.pi * sigma ** 2 / 100.0)      # integral of L times the uniform prior density
    assert abs(st["log_evidence"] - exact) < 0.25, (st["log_evidence"], exact)
    assert np.allclose(f, -0.5 * (((x - mu) / sigma) ** 2).sum(axis=1))


def test_planted_mistake_wrong_sign_is_caught():
    """A likelihood with the wrong sign puts the population at the box edge,
    far from the posterior the Gaussian test expects."""
    mu = np.array([1.0, -0.5])
    t = gh.Tmcmc([-5, -5], [5, 5], 500, max_stages=8, seed=3)
    for _ in range(40):
        x = t.candidates()
        if len(x) == 0:
            break
        t.submit(+0.5 * (((x - mu) / 0.5) ** 2).sum(axis=1))
    x, _ = t.population()
    assert np.abs(x.mean(axis=0) - mu).max() > 0.5


def test_penalty_and_errors():
    t = gh.Tmcmc([0, 0], [1, 1], 20, max_stages=3, seed=5)
    x = t.candidates()
    assert x.shape == (20, 2) and np.all((x >= 0) & (x <= 1))
    f = np.zeros(20)
    f[3] = np.nan
    f[4] = np.inf
    with pytest.raises(GHError):
        t.submit(f[:5])
    t.submit(f)
    _, pf = t.population()
    assert pf[3] == -1e12 and pf[4] == -1e12
    for _ in range(10):
        x = t.candidates()
        if len(x) == 0:
            break
        t.submit(np.zeros(len(x)))
    assert t.state()["done"]
    with pytest.raises(GHError):
        t.submit(np.zeros(1))


def test_glioma_parameter_map_matches_text_round_trip():
    rng = np.random.default_rng(9)
    lo = np.array([-8.9480, -5.9145, 6.5280, -0.9808, -0.5621, -0.5709, -4.0174, -0.5108, -0.5108, -2.9957,
                   -2.9957])
    hi = np.array([-3.2702, -1.6607, 9.8875, -0.5025, -0.2744, -0.3079, -1.3863, 0.0198, -0.2231, -0.5108,
                   -2.3026])
    x = lo + (hi - lo) * rng.uniform(size=(50, 11))
    for real in ("f32", "f64"):
        s, lk = gh.glioma_params(x, real)
        cast = (lambda v: float(np.float32(v))) if real == "f32" else float
        for i in range(len(x)):
            e = [math.exp(v) for v in x[i]]          # libm, as the library
            dw, rho, c = e[0], e[1], e[2]
            tend = math.sqrt(c / rho)
            exp_s = [cast(float("%.16f" % v)) for v in (dw, rho, tend)] + [cast(float("%.4f" % v)) for v in e[3:6]]
            exp_l = [float("%.16f" % v) for v in (e[6] ** 2, e[7], e[8], e[9], e[10])]
            assert list(s[i]) == exp_s
            assert list(lk[i]) == exp_l


def test_glioma_fitness_round_trip():
    v = np.array([1234.56789049, -3.0000004, np.nan, np.inf, 0.0])
    F = gh.glioma_fitness(v)
    assert F[0] == -1234.567890 and F[1] == -(-3.0) and F[2] == -1e12 and F[3] == -1e12 and F[4] == 0.0
