"""The FP64 tangent-linear reference pinned by central finite differences of
the oracle's own forward map (N steps, dt = T/N) and likelihood, and dJ/du
pinned by finite differences of the likelihood."""
from __future__ import annotations

import numpy as np
import pytest


def gaussian(n, x0, sigma, peak, gate):
    z, y, x = np.meshgrid(*(np.arange(n, dtype=np.float64),) * 3, indexing="ij")
    r2 = (x - x0[0]) ** 2 + (y - x0[1]) ** 2 + (z - x0[2]) ** 2
    return np.where(gate, peak * np.exp(-r2 / (2 * sigma ** 2)), 0.0)


def small_problem(seed=0):
    n = 16
    z, y, x = np.meshgrid(*(np.arange(n),) * 3, indexing="ij")
    ball = ((x - 7.5) ** 2 + (y - 7.5) ** 2 + (z - 7.5) ** 2) < 6.5 ** 2
    rng = np.random.default_rng(seed)
    pw = np.where(ball, rng.uniform(0.5, 1.0, (n, n, n)), 0.0)
    pg = np.where(ball, 1.0 - pw, 0.0)
    x0 = np.array([7.3, 8.1, 7.6])
    theta = dict(Dw=2e-4, rho=0.12, T=20.0)
    N = 10
    sigma, peak = 1.5, 0.8
    return n, pw, pg, ball, x0, theta, N, sigma, peak


def observations(oracle, n, pw, pg, ball, N):
    x_true = np.array([8.0, 7.4, 7.9])
    u0 = gaussian(n, x_true, 1.5, 0.9, ball)
    u = oracle.rd_steps(u0, pw, pg, 2.4e-4, 2.4e-5, 0.15, 1.0 / n, 25.0 / N, N)
    rng = np.random.default_rng(3)
    t1 = (u >= 0.7).astype(np.float64)
    fl = (u >= 0.25).astype(np.float64)
    for m in (t1, fl):
        f = rng.uniform(0, 1, u.shape) < 0.05
        m[f] = 1 - m[f]
    pet = np.where(t1 > 0, u + rng.normal(0, 0.05, u.shape), 0.0)
    return t1, fl, pet


def loss(oracle, n, pw, pg, ball, x0, Dw, rho, T, N, sigma, peak, obs):
    u0 = gaussian(n, x0, sigma, peak, ball)
    u = oracle.rd_steps(u0, pw, pg, Dw, 0.1 * Dw, rho, 1.0 / n, T / N, N)
    return oracle.likelihood(u, obs[2], obs[0], obs[1])[0], u


def central_fd(f, theta, k, eps, ucs=(0.7, 0.25)):
    """Central difference along axis k.  The reference likelihood jumps where a
    voxel crosses T1uc or T2uc (alpha is 1/2 just below uc and
    1/2 + 1/2(1 - e^{-1/slope}) just above, HGG_Likelihood.cpp:151-152), so
    the step is halved until no voxel changes side between theta -/+ eps."""
    for _ in range(12):
        tp, tm = theta.copy(), theta.copy()
        tp[k] += eps
        tm[k] -= eps
        (jp, up), (jm, um) = f(tp), f(tm)
        if not any(np.any((up >= uc) != (um >= uc)) for uc in ucs):
            return (jp - jm) / (2 * eps)
        eps *= 0.5
    raise AssertionError("could not find a step without threshold crossings")


@pytest.mark.parametrize("k", range(6))
def test_tlm_matches_central_fd(oracle, k):
    n, pw, pg, ball, x0, th, N, sigma, peak = small_problem()
    obs = observations(oracle, n, pw, pg, ball, N)
    theta = np.array([th["Dw"], th["rho"], th["T"], *x0])
    e = np.zeros(6)
    e[k] = 1.0
    u0 = gaussian(n, x0, sigma, peak, ball)
    J, dJ = oracle.tlm(u0, pw, pg, th["Dw"], 0.1 * th["Dw"], th["rho"], 1.0 / n, th["T"], N, sigma, x0, e,
                       obs[0], obs[1], obs[2])
    eps = 1e-4 * max(abs(theta[k]), 1.0) if k >= 3 else 1e-4 * abs(theta[k])

    def Jat(t):
        return loss(oracle, n, pw, pg, ball, t[3:], t[0], t[1], t[2], N, sigma, peak, obs)

    fd = central_fd(Jat, theta, k, eps)
    assert J == pytest.approx(Jat(theta)[0], rel=1e-13)
    assert abs(dJ) > 1e-6 * abs(J)  # a nontrivial direction
    assert dJ == pytest.approx(fd, rel=2e-5, abs=1e-8 * abs(J))


def test_dloss_du_matches_fd(oracle):
    rng = np.random.default_rng(1)
    n = 6
    u = rng.uniform(0, 1, (n, n, n))
    t1 = (rng.uniform(0, 1, u.shape) < 0.4).astype(float)
    fl = (rng.uniform(0, 1, u.shape) < 0.6).astype(float)
    pet = np.where(t1 > 0, rng.uniform(0, 1, u.shape), 0.0)
    g = oracle.dloss_du(u, t1, fl, pet)
    for c in [(0, 0, 0), (1, 2, 3), (5, 4, 1), (3, 3, 3)]:
        up, um = u.copy(), u.copy()
        up[c] += 1e-6
        um[c] -= 1e-6
        fd = (oracle.likelihood(up, pet, t1, fl)[0] - oracle.likelihood(um, pet, t1, fl)[0]) / 2e-6
        assert g[c] == pytest.approx(fd, rel=1e-5, abs=1e-6)
