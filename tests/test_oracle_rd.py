"""The oracle's RD restatement pinned against the reference.

1. golden fixtures written by tests/golden/make_golden.py from the reference's
   own ReactionDiffusionOperator/UpdateTumor (oracle/_ref): bit-exact;
2. the live reference build, when oracle/_ref exists: bit-exact on fresh inputs;
3. closed forms and invariants that a dropped term, a wrong sign or a wrong
   neighbour index would break.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

from conftest import GOLDEN, ref_available, run_rd_ref

CASES = ["rd_f32_phantom32", "rd_f64_phantom32", "rd_f32_random16", "rd_f64_random16"]


def load_case(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    return {k: z[k] for k in z.files}


@pytest.mark.parametrize("name", CASES)
def test_oracle_matches_reference_golden(oracle, name):
    c = load_case(name)
    out = oracle.rd_steps(c["phi0"], c["p_w"], c["p_g"], float(c["Dw"]), float(c["Dg"]), float(c["rho"]),
                          float(c["h"]), float(c["dt"]), int(c["nsteps"]))
    assert out.dtype == c["phi_ref"].dtype
    np.testing.assert_array_equal(out.view(np.uint8), c["phi_ref"].view(np.uint8))


@pytest.mark.skipif(not ref_available("rd_ref_f32"), reason="oracle/_ref not built (no reference tree)")
@pytest.mark.parametrize("real", ["f32", "f64"])
def test_oracle_matches_live_reference(oracle, tmp_path, real):
    t = np.float32 if real == "f32" else np.float64
    rng = np.random.default_rng(123)
    n = 16
    pw = rng.uniform(0, 1, (n, n, n)).astype(t)
    pg = (rng.uniform(0, 1, (n, n, n)) * (1 - pw)).astype(t)
    pw[rng.uniform(0, 1, pw.shape) < 0.2] = 0
    pg[rng.uniform(0, 1, pw.shape) < 0.2] = 0
    phi = (rng.uniform(0, 1, (n, n, n)) ** 4).astype(t)
    args = (1.7e-5, 1.7e-6, 0.03, 1.0 / n, 20.0, 12)
    ref = run_rd_ref(real, phi, pw, pg, *args, str(tmp_path))
    out = oracle.rd_steps(phi, pw, pg, *args)
    np.testing.assert_array_equal(out.view(np.uint8), ref.view(np.uint8))


def test_schedule_matches_reference_formula(oracle):
    """dt = cfl*h^2/(6 max(Dw,Dg)) in Real and the float step count of while(t<=tend)."""
    Dw, Dg, dt, ns = oracle.schedule("f64", 0.013, 17.6, 300.0, 64, 0.9)
    assert Dw == pytest.approx(0.013 / 17.6 ** 2, rel=1e-15)
    assert Dg == pytest.approx(0.1 * Dw, rel=1e-15)
    assert dt == pytest.approx(0.9 / 64 ** 2 / (6 * Dw), rel=1e-15)
    assert ns == int(np.floor(300.0 / dt)) + 1
    Dw32, Dg32, dt32, ns32 = oracle.schedule("f32", 0.013, 17.6, 300.0, 64, 0.9)
    t = np.float32(0)
    k = 0
    while t <= np.float32(300.0):
        t = np.float32(t + np.float32(dt32))
        k += 1
    assert k == ns32
    assert abs(ns32 - ns) <= 1


def _uniform(n, t, pw=1.0, pg=0.0):
    return np.full((n, n, n), pw, t), np.full((n, n, n), pg, t)


@pytest.mark.parametrize("real", ["f32", "f64"])
def test_constant_diffusivity_discrete_laplacian(oracle, real):
    """Constant tissue, rho=0: one step is phi + dt*D*lap_h(phi) with the
    periodic 7-point Laplacian (harmonic mean of equal values is the value)."""
    t = np.float32 if real == "f32" else np.float64
    n = 16
    pw, pg = _uniform(n, t)
    rng = np.random.default_rng(1)
    phi = (0.2 + 0.3 * rng.uniform(0, 1, (n, n, n))).astype(t)
    h, Dw, dt = 1.0 / n, 1e-4, 0.05
    out = oracle.rd_steps(phi, pw, pg, Dw, 0.1 * Dw, 0.0, h, dt, 1)
    p = phi.astype(np.float64)
    lap = sum(np.roll(p, s, a) for a in range(3) for s in (-1, 1)) - 6 * p
    expect = p + dt * float(t(Dw)) * lap / h ** 2
    tol = 1e-6 if real == "f32" else 1e-13
    np.testing.assert_allclose(out, expect, rtol=0, atol=tol)


@pytest.mark.parametrize("real", ["f32", "f64"])
def test_quadratic_field_exact_laplacian(oracle, real):
    """phi = a + b*(x-c)^2 in the interior has lap_h = 2b exactly (per axis)."""
    t = np.float32 if real == "f32" else np.float64
    n = 16
    pw, pg = _uniform(n, t)
    x = np.arange(n, dtype=np.float64)
    q = 0.1 + 1e-3 * (x - 7.5) ** 2
    phi = (q[None, None, :] + q[None, :, None] + q[:, None, None]).astype(t)
    h, Dw, dt = 1.0 / n, 1e-4, 0.01
    out = oracle.rd_steps(phi, pw, pg, Dw, 0.1 * Dw, 0.0, h, dt, 1).astype(np.float64)
    inner = (slice(1, -1),) * 3
    expect = phi.astype(np.float64) + dt * float(t(Dw)) * (3 * 2e-3) / h ** 2
    tol = 1e-6 if real == "f32" else 1e-12
    np.testing.assert_allclose(out[inner], expect[inner], rtol=0, atol=tol)


def test_logistic_recurrence_reaction_only(oracle):
    """D = 0 (no tissue diffusivity, p_w=p_g=0 would leave the domain, so use
    Dw=0 with tissue): each cell follows phi <- phi + dt*rho*phi*(1-phi), brute-forced in double."""
    n = 8
    pw, pg = _uniform(n, np.float64)
    phi = np.linspace(0.01, 0.9, n ** 3).reshape(n, n, n)
    rho, dt, k = 0.3, 0.7, 9
    out = oracle.rd_steps(phi, pw, pg, 0.0, 0.0, rho, 1.0 / n, dt, k)
    p = phi.copy()
    for _ in range(k):
        p = np.clip(p + dt * (rho * p * (1 - p)), 0, 1)
    np.testing.assert_array_equal(out, p)


def test_mass_conservation_periodic_tissue_rho0(oracle):
    """rho = 0, variable tissue everywhere (no outside cell, so no ghost-mirror
    faces): each face flux enters one cell and leaves its neighbour with the
    same harmonic coefficient, so sum(phi) is conserved to round-off."""
    n = 16
    rng = np.random.default_rng(8)
    pw = rng.uniform(0.2, 1.0, (n, n, n))
    pg = 1.0 - pw
    phi = rng.uniform(0, 0.6, (n, n, n))
    out = oracle.rd_steps(phi, pw, pg, 2e-4, 2e-5, 0.0, 1.0 / n, 0.2, 40)
    assert abs(out.sum() - phi.sum()) < 1e-12 * phi.sum()


def test_no_flux_tissue_ball_keeps_csf_empty(oracle):
    """Tissue ball inside CSF: the harmonic face diffusivity towards CSF is 0,
    so no tumour ever enters CSF.  (The ghost-mirror rule of
    ReactionDiffusionOperator.h:166-181 doubles only the boundary cell's
    inner face, so total mass is not conserved there -- reference behaviour.)"""
    n = 16
    z, y, x = np.meshgrid(*(np.arange(n),) * 3, indexing="ij")
    ball = ((x - 7.5) ** 2 + (y - 8) ** 2 + (z - 7) ** 2) < 36
    pw = np.where(ball, 0.7, 0.0)
    pg = np.where(ball, 0.3, 0.0)
    phi = np.where(ball, np.exp(-((x - 6) ** 2 + (y - 9) ** 2 + (z - 7) ** 2) / 4.0) * 0.5, 0.0)
    out = oracle.rd_steps(phi, pw, pg, 2e-4, 2e-5, 0.0, 1.0 / n, 0.2, 40)
    assert np.all(out[~ball] == 0)
    assert out.sum() > 0.9 * phi.sum()


def test_mirror_symmetry(oracle):
    """A mirror-symmetric anatomy and IC stay mirror-symmetric (checks the
    +/- neighbour pairing and the opposite-face no-flux rule)."""
    n = 16
    rng = np.random.default_rng(3)
    half = rng.uniform(0, 1, (n, n, n // 2))
    pw = np.concatenate([half, half[:, :, ::-1]], axis=2)
    pw[pw < 0.3] = 0.0
    pg = 0.5 * (1 - pw) * (pw > 0)
    ph = rng.uniform(0, 0.5, (n, n, n // 2))
    phi = np.concatenate([ph, ph[:, :, ::-1]], axis=2) * (pw > 0)
    # mirror about the plane between x=7 and x=8 maps x -> 15-x; periodic wrap maps -1 -> 15 consistently
    out = oracle.rd_steps(phi, pw, pg, 3e-4, 3e-5, 0.05, 1.0 / n, 0.1, 10)
    np.testing.assert_array_equal(out, out[:, :, ::-1])


def test_no_flux_bc_opposite_face_doubling(oracle):
    """Single tissue cell row next to CSF: the coefficient opposite an outside
    neighbour is doubled (ghost-mirror), ReactionDiffusionOperator.h:166-181."""
    n = 8
    pw = np.zeros((n, n, n))
    pw[4, 4, 2:6] = 1.0  # x = 2..5 tissue, x=1 and x=6 outside
    pg = np.zeros_like(pw)
    phi = np.zeros_like(pw)
    phi[4, 4, 2] = 0.4
    phi[4, 4, 3] = 0.1
    D, h, dt = 1.0, 1.0, 0.01
    out = oracle.rd_steps(phi, pw, pg, D, 0.1 * D, 0.0, h, dt, 1)
    # cell x=2: -x neighbour outside (chf=0<0.1) -> +x face doubled: 2*D*(0.1-0.4)
    # y/z neighbours are outside too but their faces have D=0, and each doubles a zero face
    assert out[4, 4, 2] == pytest.approx(0.4 + dt * 2 * D * (0.1 - 0.4), abs=1e-15)
    # cell x=3: both x neighbours inside (chf = 0.4+1 and 0+1): plain
    assert out[4, 4, 3] == pytest.approx(0.1 + dt * D * (0.4 - 0.1 + 0.0 - 0.1), abs=1e-15)


def test_harmonic_mean_eps_branch(oracle):
    """df_i + df_loc < 1e-8 gives a zero face, otherwise 2 d d0/(d+d0)."""
    n = 8
    pw = np.full((n, n, n), 1.0)
    pg = np.zeros_like(pw)
    pw[3, 3, 4] = 2e-9  # tiny neighbour of (3,3,3): d+d0 >= eps, harmonic ~ 2*d
    phi = np.zeros_like(pw)
    phi[3, 3, 3] = 0.5
    D, h, dt = 1.0, 1.0, 0.01
    out = oracle.rd_steps(phi, pw, pg, D, 0.0, 0.0, h, dt, 1)
    dtiny = 2e-9
    hm = 2 * dtiny * 1.0 / (dtiny + 1.0)
    # (3,3,3): the tiny neighbour is also "outside" (chf = 2e-9 < 0.1), which
    # doubles the opposite (-x) face: 4 unit faces + 2 + the harmonic face
    assert out[3, 3, 3] == pytest.approx(0.5 - dt * (4 + 2 + hm) * 0.5, abs=1e-15)
    pw2 = np.zeros_like(pw)
    pw2[3, 3, 3] = 4e-9  # d0 alone below 1e-8 with zero neighbours: every face is 0
    out2 = oracle.rd_steps(phi, pw2, pg, D, 0.0, 0.0, h, dt, 1)
    assert out2[3, 3, 3] == 0.5


def test_planted_mistakes_are_caught(oracle):
    """A restatement with a dropped reaction term or a flipped no-flux face
    differs from the reference golden vector (the checks above would fail)."""
    c = load_case("rd_f64_phantom32")
    good = oracle.rd_steps(c["phi0"], c["p_w"], c["p_g"], float(c["Dw"]), float(c["Dg"]), float(c["rho"]),
                           float(c["h"]), float(c["dt"]), int(c["nsteps"]))
    bad = oracle.rd_steps(c["phi0"], c["p_w"], c["p_g"], float(c["Dw"]), float(c["Dg"]), 0.0,
                          float(c["h"]), float(c["dt"]), int(c["nsteps"]))
    assert np.array_equal(good, c["phi_ref"])
    assert not np.array_equal(bad, c["phi_ref"])
