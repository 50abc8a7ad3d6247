"""The oracle's initial condition (Glioma_ReactionDiffusion.cpp:79-172) pinned
by what fixes it independently of its own arithmetic:

1. voxel mapping: data of the grid's own size maps cell i to voxel i; data of
   half the size repeats each voxel over 2^3 cells; data of a non-cubic,
   non-power-of-two size maps cell i to voxel floor(i*max(nx,ny,nz)/n) (exact
   rationals, up to the reference's rounding where that ratio is an integer),
   and cells past the data along each axis stay empty (catches a transposed
   axis, a wrong scale or a dropped bounds test);
2. the tissue vote equals the pinned per-voxel vote (gho_anatomy);
3. the tumour profile equals phi = (1 - psi - sin(pi psi)/pi)/2 evaluated in
   double on |x - ic| (tolerance 2e-6 for FP32 arithmetic), is 0 outside the
   smoothing shell and on cells without tissue, decreases with distance and
   is symmetric about a cell-centred seed.
"""
from __future__ import annotations

import numpy as np
import pytest

import gliomasolver_amd as gh


def _rand_anatomy(shape, seed):
    rng = np.random.default_rng(seed)
    gm, wm, csf = (rng.uniform(0, 1, shape) for _ in range(3))
    # some empty voxels and some exact ties
    gm[rng.uniform(size=shape) < 0.1] = 0
    wm[rng.uniform(size=shape) < 0.1] = 0
    csf[rng.uniform(size=shape) < 0.05] = 0
    return gm, wm, csf


@pytest.mark.parametrize("real", ["f32", "f64"])
def test_same_size_mapping_is_identity(oracle, real):
    gm, wm, csf = _rand_anatomy((32, 32, 32), 1)
    pw, pg, pc, phi = oracle.ic(4, gm, wm, csf, None, real)
    ew, eg, ec = oracle.anatomy(gm, wm, csf, real)
    assert np.array_equal(pw, ew) and np.array_equal(pg, eg) and np.array_equal(pc, ec)
    assert not phi.any()


@pytest.mark.parametrize("real", ["f32", "f64"])
def test_half_size_data_repeats_voxels(oracle, real):
    gm, wm, csf = _rand_anatomy((16, 16, 16), 2)
    pw, pg, pc, _ = oracle.ic(4, gm, wm, csf, None, real)
    ew, eg, ec = oracle.anatomy(gm, wm, csf, real)
    up = lambda a: np.repeat(np.repeat(np.repeat(a, 2, 0), 2, 1), 2, 2)
    assert np.array_equal(pw, up(ew)) and np.array_equal(pg, up(eg)) and np.array_equal(pc, up(ec))


def _exact_map(n, bmax, size):
    """voxel index of each cell, exact rationals; -1 past the data; also a
    mask of cells where i*bmax/n is an integer (reference may floor one lower)."""
    i = np.arange(n)
    num = i * bmax
    v = num // n
    tie = (num % n) == 0
    v = np.where(v < size, v, -1)
    return v, tie


@pytest.mark.parametrize("real", ["f32", "f64"])
def test_non_cubic_mapping_and_axes(oracle, real):
    nz, ny, nx = 13, 29, 21          # bmax 29 along y
    n = 32
    # voxel value encodes its own (x, y, z) index: gm = x + 100 y + 10000 z + 1
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    code = (x + 100 * y + 10000 * z + 1).astype(np.float64)
    gm = code
    wm = np.zeros_like(gm)
    csf = np.zeros_like(gm)
    pw, pg, pc, _ = oracle.ic(4, gm, wm, csf, None, real)
    # pg == 1 wherever a voxel was sampled (pure GM after normalisation), 0 elsewhere
    vx, tx = _exact_map(n, 29, nx)
    vy, ty = _exact_map(n, 29, ny)
    vz, tz = _exact_map(n, 29, nz)
    inside = (vz[:, None, None] >= 0) & (vy[None, :, None] >= 0) & (vx[None, None, :] >= 0)
    tie = tz[:, None, None] | ty[None, :, None] | tx[None, None, :]
    got = pg == 1
    assert np.array_equal(got[~tie], inside[~tie])
    assert not pw.any() and not pc.any()
    # the same cells are sampled when the tissue is white matter instead
    pw3, _, _, _ = oracle.ic(4, np.zeros_like(gm), code, csf, None, "f64")
    assert np.array_equal(pw3 == 1, pg == 1)


def test_non_cubic_mapping_picks_the_right_voxel(oracle):
    """Recover the sampled voxel index through the vote: gm = code, wm = 1 gives
    p_g = code/(code+1), which is invertible."""
    nz, ny, nx = 13, 29, 21
    n = 32
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    code = (x + 100 * y + 10000 * z + 1).astype(np.float64)
    pw, pg, _, _ = oracle.ic(4, code, np.ones_like(code), np.zeros_like(code), None, "f64")
    sampled = pw > 0
    rec = np.rint(pg[sampled] / pw[sampled]).astype(np.int64) - 1
    ci = np.argwhere(sampled)                     # (k, j, i)
    vz, vy, vx = rec // 10000, (rec // 100) % 100, rec % 100
    for (axis_v, col, size) in ((vx, 2, nx), (vy, 1, ny), (vz, 0, nz)):
        cell = ci[:, col]
        exact = (cell * 29) / n
        lo = np.floor(exact).astype(np.int64)
        ok = (axis_v == lo) | ((exact == lo) & (axis_v == lo - 1))
        assert ok.all()
        assert axis_v.max() < size


@pytest.mark.parametrize("real", ["f32", "f64"])
def test_tumour_profile_closed_form(oracle, real):
    n = 64
    bpd = n // 8
    gm = np.zeros((n, n, n))
    wm = np.ones((n, n, n))
    csf = np.zeros((n, n, n))
    wm[:, :, :20] = 0                      # tissue-free slab along x
    csf[:, :, :20] = 1
    ic = (24.0 / n, 30.0 / n, 33.0 / n)    # a cell position
    pw, pg, pc, phi = oracle.ic(bpd, gm, wm, csf, ic, real)
    h = 1.0 / n
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    d = np.sqrt((i * h - ic[0]) ** 2 + (j * h - ic[1]) ** 2 + (k * h - ic[2]) ** 2)
    psi = (d - 0.005) * 64.0
    expect = np.where(psi <= 1, 0.5 * (1 - psi - np.sin(np.pi * psi) / np.pi), 0.0)
    expect[:, :, :20] = 0
    tol = 2e-6 if real == "f32" else 1e-12
    assert np.abs(phi - expect).max() <= tol
    assert phi.max() < 1 and phi.min() >= 0
    centre = phi[33, 30, 24]
    assert abs(centre - 0.5 * (1 + 0.32 + np.sin(0.32 * np.pi) / np.pi)) < 1e-6
    # symmetric about the seed cell and decreasing outwards along each axis
    assert phi[33, 30, 25] == phi[33, 30, 23] and phi[33, 31, 24] == phi[33, 29, 24]
    assert centre > phi[33, 30, 25] > phi[33, 30, 26]
    assert not phi[d > 0.005 + 1 / 64 + 1e-9].any()


def test_phantom_ic_vote_matches_pinned_vote(oracle):
    gm, wm, csf, sv = gh.phantom(32, 5)
    for real in ("f32", "f64"):
        pw, pg, pc, _ = oracle.ic(4, gm, wm, csf, None, real)
        ew, eg, ec = oracle.anatomy(gm, wm, csf, real)
        assert np.array_equal(pw, ew) and np.array_equal(pg, eg) and np.array_equal(pc, ec)


def test_planted_mistake_transposed_axes_is_caught(oracle):
    """Swapping x and z in the data is visible to the mapping test's checks."""
    nz, ny, nx = 13, 29, 21
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    code = (x + 100 * y + 10000 * z + 1).astype(np.float64)
    pw, pg, _, _ = oracle.ic(4, np.ascontiguousarray(code.transpose(2, 1, 0)),
                             np.ones((nx, ny, nz)), np.zeros((nx, ny, nz)), None, "f64")
    assert not np.array_equal(pw > 0, oracle.ic(4, code, np.ones_like(code), np.zeros_like(code), None, "f64")[0] > 0)
