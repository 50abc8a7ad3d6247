"""GPU parity of the HIP reaction-diffusion path (through the C ABI) against
the reference's golden vectors and the oracle.  Bar: bit-exact for both Real
kinds (the kernels reproduce the reference's rounding points)."""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

import gliomasolver_amd as gh
from gliomasolver_amd import FIELD_DPHIDT, FIELD_PG, FIELD_PHI, FIELD_PW
from gliomasolver_amd import workloads as wl
from conftest import GOLDEN

pytestmark = pytest.mark.gpu

CASES = ["rd_f32_phantom32", "rd_f64_phantom32", "rd_f32_random16", "rd_f64_random16"]


def say(*a):
    print("[gpu]", *a, flush=True)
    sys.stdout.flush()


def load_case(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    return {k: z[k] for k in z.files}


def gpu_run(real, phi0, pw, pg, Dw, Dg, rho, dt, nsteps, two_pass=False):
    n = phi0.shape[0]
    ctx = gh.Context(0, real)
    ctx.grid_init_uniform(n // 8)
    ctx.set_field(FIELD_PW, pw)
    ctx.set_field(FIELD_PG, pg)
    ctx.set_field(FIELD_PHI, phi0)
    ctx.rd_prepare(Dw, Dg, rho)
    if two_pass:
        for _ in range(nsteps):
            ctx.rd_rhs()
            ctx.rd_update(dt)
    else:
        ctx.rd_steps(nsteps, dt)
    out = ctx.get_field(FIELD_PHI)
    ctx.close()
    return out


def bits_equal(a, b):
    return a.dtype == b.dtype and np.array_equal(a.view(np.uint8), b.view(np.uint8))


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("two_pass", [False, True])
def test_gpu_matches_reference_golden(name, two_pass):
    c = load_case(name)
    real = "f32" if c["phi0"].dtype == np.float32 else "f64"
    say(name, "two_pass" if two_pass else "fused", "steps", int(c["nsteps"]))
    out = gpu_run(real, c["phi0"], c["p_w"], c["p_g"], float(c["Dw"]), float(c["Dg"]), float(c["rho"]),
                  float(c["dt"]), int(c["nsteps"]), two_pass)
    diff = np.abs(out.astype(np.float64) - c["phi_ref"].astype(np.float64))
    assert bits_equal(out, c["phi_ref"]), "max |diff| %g at %s" % (diff.max(), np.unravel_index(diff.argmax(),
                                                                                                  diff.shape))


def test_gpu_rhs_matches_oracle_first_step(oracle):
    """dphidt of the two-pass API equals the oracle's after one step (phi1 = phi0 + dt*dphidt)."""
    c = load_case("rd_f64_random16")
    say("rhs f64 random16")
    n = c["phi0"].shape[0]
    ctx = gh.Context(0, "f64")
    ctx.grid_init_uniform(n // 8)
    ctx.set_field(FIELD_PW, c["p_w"])
    ctx.set_field(FIELD_PG, c["p_g"])
    ctx.set_field(FIELD_PHI, c["phi0"])
    ctx.rd_prepare(float(c["Dw"]), float(c["Dg"]), float(c["rho"]))
    ctx.rd_rhs()
    d = ctx.get_field(FIELD_DPHIDT)
    ctx.close()
    one = oracle.rd_steps(c["phi0"], c["p_w"], c["p_g"], float(c["Dw"]), float(c["Dg"]), float(c["rho"]),
                          float(c["h"]), 1.0, 1)
    # with dt = 1: phi1 = clamp(phi0 + dphidt)
    expect = np.clip(c["phi0"] + d, 0, 1)
    assert np.array_equal(one, expect)


def test_gpu_uniform64_fp64_full_run_bit_exact(oracle):
    """64^3 FP64, Dw=0.013, rho=0.025, T=300 d, CFL 0.9 (whole run) vs the oracle."""
    case = wl.config_uniform64()
    say("uniform64 f64 full run", case.describe())
    ctx = wl.load_context(case)
    pw, pg, phi0 = ctx.get_field(FIELD_PW), ctx.get_field(FIELD_PG), ctx.get_field(FIELD_PHI)
    ctx.rd_steps(case.nsteps, case.dt)
    out = ctx.get_field(FIELD_PHI)
    ctx.close()
    opw, opg, _ = oracle.anatomy(case.gm, case.wm, case.csf, "f64")
    assert bits_equal(pw, opw) and bits_equal(pg, opg)
    ref = oracle.rd_steps(phi0, pw, pg, case.Dw, case.Dg, case.rho, 1.0 / case.n, case.dt, case.nsteps)
    assert bits_equal(out, ref)
    assert out.max() > 0.1  # the tumour grew


def test_gpu_uniform128_fp32_vs_oracle_and_fp64_error(oracle):
    """128^3 FP32 (log-uniform Dw, rho draw): 60 steps bit-exact vs the FP32
    oracle; error of the FP32 state against the FP64 oracle stated and < 1e-3 rel-L2."""
    case = wl.config_uniform128()
    k = 60
    say("uniform128 f32", k, "steps", case.describe())
    ctx = wl.load_context(case)
    pw, pg, phi0 = ctx.get_field(FIELD_PW), ctx.get_field(FIELD_PG), ctx.get_field(FIELD_PHI)
    ctx.rd_steps(k, case.dt)
    out = ctx.get_field(FIELD_PHI)
    ctx.close()
    ref = oracle.rd_steps(phi0, pw, pg, case.Dw, case.Dg, case.rho, 1.0 / case.n, case.dt, k)
    assert bits_equal(out, ref)
    pw64, pg64, _ = oracle.anatomy(case.gm, case.wm, case.csf, "f64")
    Dw, Dg, dt, _ = gh.schedule("f64", case.Dw_phys, wl.L_CUBE, case.tend, case.n, case.cfl)
    ref64 = oracle.rd_steps(phi0.astype(np.float64), pw64, pg64, Dw, Dg, case.rho, 1.0 / case.n, dt, k)
    rel = np.linalg.norm(out - ref64) / np.linalg.norm(ref64)
    say("fp32 state vs fp64 oracle after %d steps: rel-L2 %.3e, max-abs %.3e" % (k, rel, np.abs(out - ref64).max()))
    assert rel < 1e-3


def test_gpu_checkpoint_restart_is_exact():
    c = load_case("rd_f32_phantom32")
    n = c["phi0"].shape[0]
    ctx = gh.Context(0, "f32")
    ctx.grid_init_uniform(n // 8)
    ctx.set_field(FIELD_PW, c["p_w"])
    ctx.set_field(FIELD_PG, c["p_g"])
    ctx.set_field(FIELD_PHI, c["phi0"])
    ctx.rd_prepare(float(c["Dw"]), float(c["Dg"]), float(c["rho"]))
    ctx.checkpoint_save(0)
    ctx.rd_steps(int(c["nsteps"]), float(c["dt"]))
    a = ctx.get_field(FIELD_PHI)
    ctx.checkpoint_load(0)
    ctx.rd_steps(int(c["nsteps"]), float(c["dt"]))
    b = ctx.get_field(FIELD_PHI)
    ctx.close()
    assert bits_equal(a, b) and bits_equal(a, c["phi_ref"])


@pytest.mark.parametrize("name", ["rd_f32_random16", "rd_f64_phantom32"])
@pytest.mark.parametrize("bpw", [0, 1, 2, 4])
def test_gpu_tiles_are_bit_identical(bpw, name):
    """Every schedule (0: two Euler steps per launch) equals the reference's golden result."""
    c = load_case(name)
    say("tile", bpw, name)
    n = c["phi0"].shape[0]
    ctx = gh.Context(0, "f32" if "f32" in name else "f64")
    ctx.grid_init_uniform(n // 8)
    ctx.set_field(FIELD_PW, c["p_w"])
    ctx.set_field(FIELD_PG, c["p_g"])
    ctx.set_field(FIELD_PHI, c["phi0"])
    ctx.rd_prepare(float(c["Dw"]), float(c["Dg"]), float(c["rho"]))
    ctx.rd_set_tile(bpw)
    ctx.rd_steps(int(c["nsteps"]), float(c["dt"]))
    out = ctx.get_field(FIELD_PHI)
    ctx.close()
    assert bits_equal(out, c["phi_ref"])


def test_gpu_autotune_restores_state():
    c = load_case("rd_f64_phantom32")
    say("autotune f64 phantom32")
    n = c["phi0"].shape[0]
    ctx = gh.Context(0, "f64")
    ctx.grid_init_uniform(n // 8)
    ctx.set_field(FIELD_PW, c["p_w"])
    ctx.set_field(FIELD_PG, c["p_g"])
    ctx.set_field(FIELD_PHI, c["phi0"])
    ctx.rd_prepare(float(c["Dw"]), float(c["Dg"]), float(c["rho"]))
    best = ctx.rd_autotune(float(c["dt"]), 5)
    assert best in (0, 1, 2, 4)
    assert bits_equal(ctx.get_field(FIELD_PHI), c["phi0"])
    ctx.rd_steps(int(c["nsteps"]), float(c["dt"]))
    out = ctx.get_field(FIELD_PHI)
    ctx.close()
    assert bits_equal(out, c["phi_ref"])


@pytest.mark.parametrize("n,nsteps", [(8, 7), (16, 9), (32, 12), (64, 1)])
def test_gpu_two_step_schedule_matches_oracle(oracle, n, nsteps):
    """Two steps per launch, with an odd remainder step, against the oracle on
    grids of 1, 2, 4 and 8 blocks per side; bpd 1 makes a block its own
    neighbour in every direction."""
    rng = np.random.default_rng(n + nsteps)
    pw = rng.uniform(0, 1, (n, n, n)).astype(np.float32)
    pg = (rng.uniform(0, 1, (n, n, n)) * (1 - pw)).astype(np.float32)
    pw[rng.uniform(size=pw.shape) < 0.15] = 0
    pg[rng.uniform(size=pg.shape) < 0.15] = 0
    phi = rng.uniform(0, 1, (n, n, n)).astype(np.float32)
    say("two-step schedule n=%d steps=%d" % (n, nsteps))
    ctx = gh.Context(0, "f32")
    ctx.grid_init_uniform(n // 8)
    ctx.set_field(FIELD_PW, pw)
    ctx.set_field(FIELD_PG, pg)
    ctx.set_field(FIELD_PHI, phi)
    Dw, Dg, rho, dt = 0.0021, 0.00021, 0.31, 0.23 / (n * n) / 0.0021
    ctx.rd_prepare(Dw, Dg, rho)
    ctx.rd_set_tile(0)
    ctx.rd_steps(nsteps, dt)
    out = ctx.get_field(FIELD_PHI)
    ctx.close()
    ref = oracle.rd_steps(phi, pw, pg, np.float32(Dw), np.float32(Dg), np.float32(rho), np.float32(1.0 / n), dt,
                          nsteps)
    assert bits_equal(out, ref)


@pytest.mark.parametrize("real", ["f32", "f64"])
def test_gpu_clamped_pairs_match_oracle_on_edge_states(oracle, real):
    """Two steps per launch on clamped states (sign-encoded outside flags and
    one constant record per cell) against the oracle: an initial state with
    negative values, -0 and values above 1, cells with p_w + p_g = 0 (domain
    boundary), = 0.1 - phi (outside threshold) and >= 0.1 (never outside)."""
    n, nsteps = 32, 11
    dt_ = np.float32 if real == "f32" else np.float64
    rng = np.random.default_rng(77)
    pw = rng.uniform(0, 1, (n, n, n))
    pg = rng.uniform(0, 1, (n, n, n)) * (1 - pw)
    z = rng.uniform(size=pw.shape)
    pw[z < 0.2] = 0
    pg[z < 0.2] = 0
    pw[(z >= 0.2) & (z < 0.3)] = 0.05
    pg[(z >= 0.2) & (z < 0.3)] = 0.0
    pw, pg = pw.astype(dt_), pg.astype(dt_)
    phi = rng.uniform(-0.2, 1.2, (n, n, n)).astype(dt_)
    phi[rng.uniform(size=phi.shape) < 0.05] = dt_(-0.0)
    phi[(z >= 0.2) & (z < 0.3)] = dt_(0.05)
    say("clamped pairs", real, "n=%d steps=%d" % (n, nsteps))
    ctx = gh.Context(0, real)
    ctx.grid_init_uniform(n // 8)
    ctx.set_field(FIELD_PW, pw)
    ctx.set_field(FIELD_PG, pg)
    ctx.set_field(FIELD_PHI, phi)
    Dw, Dg, rho = 0.0021, 0.00021, 0.31
    dt = 0.2 / (n * n) / Dw
    ctx.rd_prepare(Dw, Dg, rho)
    ctx.rd_set_tile(0)
    ctx.rd_steps(nsteps, dt)
    out = ctx.get_field(FIELD_PHI)
    ctx.close()
    ref = oracle.rd_steps(phi, pw, pg, dt_(Dw), dt_(Dg), dt_(rho), dt_(1.0 / n), dt, nsteps)
    assert bits_equal(out, ref)


def test_gpu_schedule_info_reports_the_clamped_pair_kernel():
    """p_w + p_g >= 0 everywhere enables the clamped-state pair kernel; one
    cell with p_w + p_g < 0 disables it (and results stay golden-exact)."""
    c = load_case("rd_f32_random16")
    say("schedule info f32 random16")
    n = c["phi0"].shape[0]
    pw = c["p_w"].copy()
    for negative, expect in ((False, True), (True, False)):
        p = pw.copy()
        if negative:
            p[3, 4, 5] = np.float32(-0.25)
            p_g = c["p_g"].copy()
            p_g[3, 4, 5] = np.float32(0.0)
        else:
            p = np.maximum(p, np.float32(0.0))
            p_g = c["p_g"]
        ctx = gh.Context(0, "f32")
        ctx.grid_init_uniform(n // 8)
        ctx.set_field(FIELD_PW, p)
        ctx.set_field(FIELD_PG, p_g)
        ctx.set_field(FIELD_PHI, c["phi0"])
        ctx.rd_prepare(float(c["Dw"]), float(c["Dg"]), float(c["rho"]))
        ctx.rd_set_tile(0)
        assert ctx.rd_schedule_info() == (0, expect)
        ctx.close()


@pytest.mark.parametrize("real", ["f32", "f64"])
def test_gpu_zero_flux_blocks_match_oracle(oracle, real):
    """Blocks without tissue (every face coefficient +0) take the two-step
    kernel's reaction-only path: the count equals the tissue-free blocks, and
    the state, tumour in those blocks included (phi > 0 there grows by the
    logistic term alone), equals the oracle bit for bit.  Blocks that hold an
    adjacent pair of tissue cells keep the full stencil."""
    n, nsteps = 32, 11
    dt_ = np.float32 if real == "f32" else np.float64
    rng = np.random.default_rng(5)
    pw = rng.uniform(0, 1, (n, n, n))
    pg = rng.uniform(0, 1, (n, n, n)) * (1 - pw)
    dead = 0
    for bz in range(4):
        for by in range(4):
            for bx in range(4):
                if (bx + 2 * by + bz) % 3:
                    continue
                sl = np.s_[8 * bz:8 * bz + 8, 8 * by:8 * by + 8, 8 * bx:8 * bx + 8]
                pw[sl] = 0
                pg[sl] = 0
                if (bx + by) % 2:  # a tissue pair inside: not zero-flux
                    pw[8 * bz + 3, 8 * by + 4, 8 * bx + 2:8 * bx + 4] = 0.6
                else:
                    dead += 1
    pw, pg = pw.astype(dt_), pg.astype(dt_)
    phi = rng.uniform(0, 1, (n, n, n)).astype(dt_)
    phi[rng.uniform(size=phi.shape) < 0.3] = 0
    say("zero-flux blocks", real, "expect", dead)
    ctx = gh.Context(0, real)
    ctx.grid_init_uniform(n // 8)
    ctx.set_field(FIELD_PW, pw)
    ctx.set_field(FIELD_PG, pg)
    ctx.set_field(FIELD_PHI, phi)
    Dw, Dg, rho = 0.0021, 0.00021, 0.31
    dt = 0.2 / (n * n) / Dw
    ctx.rd_prepare(Dw, Dg, rho)
    ctx.rd_set_tile(0)
    assert ctx.rd_schedule_info() == (0, True)
    assert ctx.rd_zero_flux_blocks() == dead > 0
    ctx.rd_steps(nsteps, dt)
    out = ctx.get_field(FIELD_PHI)
    ctx.close()
    ref = oracle.rd_steps(phi, pw, pg, dt_(Dw), dt_(Dg), dt_(rho), dt_(1.0 / n), dt, nsteps)
    assert bits_equal(out, ref)
    assert not np.array_equal(out[pw + pg == 0], phi[pw + pg == 0])  # the reaction did act there
