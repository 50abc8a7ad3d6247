"""GPU adjoint (through the C ABI) against the FP64 tangent-linear oracle,
central finite differences of GPU forward solves, and torch gradcheck."""
from __future__ import annotations

import sys

import numpy as np
import pytest

import gliomasolver_amd as gh
from gliomasolver_amd import FIELD_PG, FIELD_PHI, FIELD_PW
from test_oracle_tlm import central_fd, gaussian, observations, small_problem

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu]", *a, flush=True)
    sys.stdout.flush()


def gpu_setup(real, n, pw, pg, obs):
    ctx = gh.Context(0, real)
    ctx.grid_init_uniform(n // 8)
    ctx.set_field(FIELD_PW, pw)
    ctx.set_field(FIELD_PG, pg)
    ctx.lik_set_data(obs[0], obs[1], obs[2])
    return ctx


def gpu_adjoint(ctx, u0, Dw, rho, T, N, x0, sigma, K):
    ctx.set_field(FIELD_PHI, u0)
    ctx.rd_prepare(Dw, 0.1 * Dw, rho)
    return ctx.rd_adjoint(N, T, x0, sigma, K)


@pytest.mark.parametrize("real,rel", [("f64", 1e-9), ("f32", 2e-3)])
def test_gpu_adjoint_matches_tlm(oracle, real, rel):
    n, pw, pg, ball, x0, th, N, sigma, peak = small_problem()
    obs = observations(oracle, n, pw, pg, ball, N)
    u0 = gaussian(n, x0, sigma, peak, ball)
    say("adjoint vs tlm", real)
    ctx = gpu_setup(real, n, pw, pg, obs)
    loss, grad = gpu_adjoint(ctx, u0, th["Dw"], th["rho"], th["T"], N, x0, sigma, 3)
    ctx.close()
    for k in range(6):
        e = np.zeros(6)
        e[k] = 1
        J, dJ = oracle.tlm(u0, pw, pg, th["Dw"], 0.1 * th["Dw"], th["rho"], 1.0 / n, th["T"], N, sigma, x0, e,
                           obs[0], obs[1], obs[2])
        if k == 0:
            assert loss[0] == pytest.approx(J, rel=1e-12 if real == "f64" else 1e-5)
        say("  d%d adjoint %.10e tlm %.10e" % (k, grad[k], dJ))
        assert grad[k] == pytest.approx(dJ, rel=rel, abs=1e-9 * abs(J))


def test_gpu_adjoint_checkpoint_interval_invariant(oracle):
    n, pw, pg, ball, x0, th, N, sigma, peak = small_problem()
    obs = observations(oracle, n, pw, pg, ball, N)
    u0 = gaussian(n, x0, sigma, peak, ball)
    ctx = gpu_setup("f64", n, pw, pg, obs)
    ref = None
    for K in (1, 3, 4, 64):
        say("adjoint ckpt_every", K)
        loss, grad = gpu_adjoint(ctx, u0, th["Dw"], th["rho"], th["T"], N, x0, sigma, K)
        if ref is None:
            ref = (loss, grad)
        else:
            assert np.array_equal(loss, ref[0]) and np.array_equal(grad, ref[1])
    ctx.close()


def test_gpu_adjoint_matches_central_fd_of_gpu_forward(oracle):
    """FD step 1e-4 relative (shrunk if a voxel would cross T1uc/T2uc)."""
    n, pw, pg, ball, x0, th, N, sigma, peak = small_problem()
    obs = observations(oracle, n, pw, pg, ball, N)
    ctx = gpu_setup("f64", n, pw, pg, obs)
    theta = np.array([th["Dw"], th["rho"], th["T"], *x0])

    def forward(t):
        ctx.set_field(FIELD_PHI, gaussian(n, t[3:], sigma, peak, ball))
        ctx.rd_prepare(t[0], 0.1 * t[0], t[1])
        ctx.rd_steps(N, t[2] / N)
        return ctx.likelihood()[0], ctx.get_field(FIELD_PHI)

    _, grad = gpu_adjoint(ctx, gaussian(n, x0, sigma, peak, ball), th["Dw"], th["rho"], th["T"], N, x0, sigma, 4)
    for k in range(6):
        eps = 1e-4 * (abs(theta[k]) if k < 3 else max(abs(theta[k]), 1.0))
        fd = central_fd(forward, theta, k, eps)
        say("  d%d adjoint %.8e fd %.8e" % (k, grad[k], fd))
        assert grad[k] == pytest.approx(fd, rel=3e-5)
    ctx.close()


def test_gpu_adjoint_phantom64_fp32_vs_tlm(oracle):
    """64^3 phantom, FP32 state, FP64 adjoint accumulation, 64-step checkpoints;
    error against the FP64 tangent-linear oracle stated."""
    from gliomasolver_amd import workloads as wl
    case = wl.make_case(64, "f32", 4, Dw_phys=0.013, rho=0.05, tend=120.0)
    N = case.nsteps
    say("phantom64 f32 adjoint N=%d" % N)
    ctx = wl.load_context(case)
    pw, pg, u0 = ctx.get_field(FIELD_PW), ctx.get_field(FIELD_PG), ctx.get_field(FIELD_PHI)
    rng = np.random.default_rng(1)
    u_true = oracle.rd_steps(u0.astype(np.float64), pw.astype(np.float64), pg.astype(np.float64), case.Dw * 1.3,
                             case.Dg * 1.3, 0.08, 1.0 / 64, case.tend / N, N)
    t1 = (u_true >= 0.7).astype(float)
    fl = (u_true >= 0.25).astype(float)
    pet = np.where(t1 > 0, u_true + rng.normal(0, 0.05, u_true.shape), 0.0)
    ctx.lik_set_data(t1, fl, pet)
    loss, grad = ctx.rd_adjoint(N, case.tend, case.seed_voxel, 1.0, 64)
    ctx.close()
    pw64, pg64, _ = oracle.anatomy(case.gm, case.wm, case.csf, "f64")
    u064 = gh.gaussian_ic("f64", pw64, pg64, case.seed_voxel, 1.0, 0.1)
    Dw64 = 0.013 / 17.6 ** 2
    tlm = np.zeros(6)
    for k in range(6):
        e = np.zeros(6)
        e[k] = 1
        J, tlm[k] = oracle.tlm(u064, pw64, pg64, Dw64, 0.1 * Dw64, case.rho, 1.0 / 64, case.tend, N, 1.0,
                               np.array(case.seed_voxel, float), e, t1, fl, pet)
    # parameters: relative to each derivative; seed position (one unit, cells):
    # relative to the largest of the three position derivatives
    pos = np.abs(tlm[3:]).max()
    for k in range(6):
        scale = abs(tlm[k]) if k < 3 else pos
        err = abs(grad[k] - tlm[k]) / scale
        say("  d%d fp32-state adjoint %.6e fp64 tlm %.6e err %.2e" % (k, grad[k], tlm[k], err))
        assert err < 5e-3


@pytest.mark.parametrize("real", ["f32", "f64"])
def test_gpu_adjoint_tissue_free_shortcut_is_bit_identical(real):
    """Under the two-step schedule the reverse step advances blocks without
    diffusion (no tissue: every face coefficient +0) by the cell's own update,
    with no halo; the loss and all six gradients equal, bit for bit, the
    adjoint under a single-step schedule, which takes every block through the
    full transposed stencil."""
    from gliomasolver_amd import workloads as wl
    case = wl.make_case(64, real, 4, Dw_phys=0.013, rho=0.05, tend=60.0)
    N = case.nsteps
    ctx = wl.load_context(case)
    ctx.checkpoint_save(0)
    ctx.rd_steps(N, case.dt)
    u = ctx.get_field(FIELD_PHI).astype(np.float64)
    rng = np.random.default_rng(3)
    t1 = (u >= 0.5).astype(np.float64)
    fl = (u >= 0.2).astype(np.float64)
    ctx.lik_set_data(t1, fl, np.where(t1 > 0, u + rng.normal(0, 0.05, u.shape), 0.0))
    assert ctx.rd_zero_flux_blocks() > 100
    out = {}
    for tile in (0, 2):
        ctx.rd_set_tile(tile)
        ctx.checkpoint_load(0)
        out[tile] = ctx.rd_adjoint(N, case.tend, case.seed_voxel, 1.0, 16)
    ctx.close()
    say("adjoint", real, "tile 0 vs 2: loss", out[0][0][0], out[2][0][0])
    assert np.array_equal(out[0][0], out[2][0]) and np.array_equal(out[0][1], out[2][1])


def test_gpu_adjoint_uniform128_fp32_agrees_with_fp64():
    """The gradient configuration at full size: 128^3 phantom, T = 400 d, FP32
    state with FP64 adjoint and gradient sums, checkpoints every 64 steps,
    against the same solve with FP64 state; all six gradients and the loss.
    Observations from an FP64 solve with other parameters (T1Gd u >= 0.7,
    FLAIR u >= 0.25, 5 % flips, PET = u + N(0, 0.05^2) inside T1Gd)."""
    from gliomasolver_amd import workloads as wl
    out = {}
    obs = None
    N = T = None
    for real in ("f64", "f32"):
        case = wl.make_case(128, real, 2)
        ctx = wl.load_context(case)
        if obs is None:
            ctx.rd_prepare(case.Dw * 1.2, case.Dg * 1.2, case.rho * 0.9)
            ctx.rd_steps(case.nsteps, case.dt)
            u = ctx.get_field(FIELD_PHI)
            rng = np.random.default_rng(41)
            t1 = (u >= 0.7).astype(np.float64)
            fl = (u >= 0.25).astype(np.float64)
            for m in (t1, fl):
                f = rng.uniform(size=m.shape) < 0.05
                m[f] = 1 - m[f]
            obs = (t1, fl, np.where(t1 > 0, u + rng.normal(0, 0.05, u.shape), 0.0))
            ctx.close()
            ctx = wl.load_context(case)
        if N is None:  # one step count and horizon for both (the FP64 schedule's)
            N, T = case.nsteps, case.nsteps * case.dt
        ctx.lik_set_data(*obs)
        ctx.rd_prepare(case.Dw, case.Dg, case.rho)
        say("adjoint 128^3", real, "N=%d" % N)
        out[real] = ctx.rd_adjoint(N, T, case.seed_voxel, 1.0, 64)
        ctx.close()
    (l64, g64), (l32, g32) = out["f64"], out["f32"]
    lerr = abs(l32[0] - l64[0]) / abs(l64[0])
    pos = np.abs(g64[3:]).max()
    errs = [abs(g32[k] - g64[k]) / (abs(g64[k]) if k < 3 else pos) for k in range(6)]
    say("  loss f32 %.9e f64 %.9e rel %.2e" % (l32[0], l64[0], lerr))
    for k in range(6):
        say("  d%d f32 %.6e f64 %.6e err %.2e" % (k, g32[k], g64[k], errs[k]))
    assert lerr < 1e-3
    assert max(errs) < 2e-2


def test_torch_gradcheck_over_c_abi():
    torch = pytest.importorskip("torch")
    from gliomasolver_amd.autograd import AdjointProblem, neg_log_likelihood
    n = 16
    z, y, x = np.meshgrid(*(np.arange(n),) * 3, indexing="ij")
    ball = ((x - 7.5) ** 2 + (y - 7.5) ** 2 + (z - 7.5) ** 2) < 6.5 ** 2
    pw = np.where(ball, 0.8, 0.0)
    pg = np.where(ball, 0.2, 0.0)
    rng = np.random.default_rng(0)
    t1 = (rng.uniform(0, 1, pw.shape) < 0.1).astype(float)
    fl = (rng.uniform(0, 1, pw.shape) < 0.3).astype(float)
    pet = np.where(t1 > 0, rng.uniform(0, 0.5, pw.shape), 0.0)
    ctx = gh.Context(0, "f64")
    ctx.grid_init_uniform(2)
    ctx.set_field(FIELD_PW, pw)
    ctx.set_field(FIELD_PG, pg)
    ctx.lik_set_data(t1, fl, pet)
    # thresholds above every simulated value keep J smooth for torch's FD
    prob = AdjointProblem(ctx, L=17.6, nsteps=8, sigma=1.5, peak=0.4, ckpt_every=3,
                          lik=(0.000361, 0.85, 0.95, 0.9, 2.0))
    theta = torch.tensor([0.05, 0.1, 16.0, 7.3, 8.1, 7.6], dtype=torch.float64, requires_grad=True)
    say("gradcheck")
    assert torch.autograd.gradcheck(lambda t: neg_log_likelihood(t, prob), (theta,), eps=1e-6, atol=1e-4,
                                    rtol=1e-4)
    ctx.close()
