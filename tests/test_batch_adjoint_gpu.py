"""GPU parity of the batched forward + adjoint (gh_batch_adjoint).

A mixed batch (grids of 32^3, 24^3 and 16^3 cells; different Dw, rho, T,
Gaussian centres and step counts, checkpoints every 16 steps so that segments
end inside some samples' runs) must give, per sample,
  * exactly the loss and the six gradients of the single-solve adjoint
    (gh_rd_adjoint through gh_ctx) for the power-of-two grids, and
  * in FP64, the FP64 tangent-linear oracle's directional derivatives
    (1e-9 relative) for the 24^3 grid the single-solve path cannot build.
"""
from __future__ import annotations

import numpy as np
import pytest

import gliomasolver_amd as gh
from gliomasolver_amd import FIELD_PG, FIELD_PHI, FIELD_PW
from test_batch_gpu import ANAT, LIK, anatomy, observations, say

pytestmark = pytest.mark.gpu
SIGMA, PEAK, K = 1.0, 0.1, 16
SAMPLES = [  # anatomy, Dw (/L^2), rho, T, x0, y0, z0 (cells), N -- dt = T/N below 0.9 h^2 / (6 Dw)
    (0, 0.0005, 0.2, 10.0, 14.3, 16.2, 13.7, 37),
    (1, 0.0006, 0.1, 20.0, 12.0, 11.5, 12.4, 64),
    (0, 0.0008, 0.3, 15.0, 15.1, 15.9, 16.4, 100),
    (2, 0.0010, 0.4, 10.0, 8.2, 7.6, 8.1, 23),
    (1, 0.0009, 0.05, 12.0, 11.3, 12.6, 11.8, 50),
    (0, 0.0002, 0.25, 8.0, 13.0, 17.0, 14.0, 16),
]


@pytest.fixture(scope="module", params=["f32", "f64"])
def batch_adjoint(request):
    real = request.param
    say("batched adjoint", real, "%d samples, K=%d" % (len(SAMPLES), K))
    b = gh.Batch(0, real)
    data = []
    for bpd, shape, seed in ANAT:
        gm, wm, csf = anatomy(shape, seed)
        aid = b.add_anatomy(bpd, gm, wm, csf)
        obs = observations(8 * bpd, seed + 100)
        b.set_observations(aid, *obs)
        data.append((bpd, gm, wm, csf, obs))
    b.set_adjoint_samples([s[0] for s in SAMPLES], [s[1:7] for s in SAMPLES], [s[7] for s in SAMPLES], SIGMA, PEAK)
    loss, grad, ms = b.adjoint(LIK, ckpt_every=K, stepPET=2, stepTi=1)
    b.close()
    say("batched adjoint device time %.2f ms" % ms)
    return real, data, loss, grad


def test_batch_adjoint_matches_single_solve(batch_adjoint):
    real, data, loss, grad = batch_adjoint
    for i, (a, Dw, rho, T, x0, y0, z0, N) in enumerate(SAMPLES):
        bpd, gm, wm, csf, obs = data[a]
        if bpd & (bpd - 1):
            continue
        with gh.Context(0, real) as ctx:
            ctx.grid_init_uniform(bpd)
            ctx.ic_from_anatomy(gm, wm, csf, None)
            pw, pg = ctx.get_field(FIELD_PW), ctx.get_field(FIELD_PG)
            ctx.set_field(FIELD_PHI, gh.gaussian_ic(real, pw, pg, (x0, y0, z0), SIGMA, PEAK))
            ctx.rd_prepare(Dw, Dw / 10.0, rho)
            ctx.lik_set_data(*obs)
            l1, g1 = ctx.rd_adjoint(N, T, (x0, y0, z0), SIGMA, K, params=LIK, stepPET=2, stepTi=1)
        say("sample %d loss %.12e grad" % (i, l1[0]), np.array2string(g1, precision=6))
        assert np.array_equal(loss[i], l1), "sample %d" % i
        assert np.array_equal(grad[i], g1), "sample %d" % i


def test_batch_adjoint_matches_tlm_oracle(batch_adjoint, oracle):
    real, data, loss, grad = batch_adjoint
    if real != "f64":
        pytest.skip("the 1e-9 tangent-linear comparison is for the FP64 state")
    for i, (a, Dw, rho, T, x0, y0, z0, N) in enumerate(SAMPLES):
        bpd, gm, wm, csf, obs = data[a]
        if not bpd & (bpd - 1):
            continue
        n = 8 * bpd
        pw, pg, _, _ = oracle.ic(bpd, gm, wm, csf, None, "f64")
        u0 = gh.gaussian_ic("f64", pw, pg, (x0, y0, z0), SIGMA, PEAK)
        for k in range(6):
            e = np.zeros(6)
            e[k] = 1
            J, dJ = oracle.tlm(u0, pw, pg, Dw, 0.1 * Dw, rho, 1.0 / n, T, N, SIGMA, (x0, y0, z0), e, obs[0], obs[1],
                               obs[2], lik=LIK, stepPET=2, stepTi=1)
            if k == 0:
                assert loss[i, 0] == pytest.approx(J, rel=1e-12)
            say("  sample %d d%d batch %.10e tlm %.10e" % (i, k, grad[i, k], dJ))
            assert grad[i, k] == pytest.approx(dJ, rel=1e-9, abs=1e-9 * abs(J))
