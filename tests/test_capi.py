"""C ABI checks that need no GPU: the library loads, exports every entry point
declared in include/gliomahip.h, validates arguments, and its host-only
functions (schedule, phantom, Gaussian IC) agree with the oracle."""
from __future__ import annotations

import ctypes as C
import subprocess

import numpy as np
import pytest

import gliomasolver_amd as gh
from gliomasolver_amd import _capi


def test_library_exports_every_declared_symbol():
    lib = _capi.load()
    names = _capi.declared_symbols()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", _capi.LIB_PATH], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    assert set(names) <= exported
    # every declared symbol has a ctypes signature in the binding
    assert set(names) <= set(_capi.SIGNATURES)


def test_argument_errors_do_not_touch_the_device():
    lib = _capi.load()
    h = C.c_void_p()
    assert lib.gh_ctx_create(0, 7, C.byref(h)) == -1
    assert b"real_kind" in lib.gh_last_error()
    d = C.c_double()
    n = C.c_int64()
    assert lib.gh_schedule(0, 0.013, 17.6, 300.0, 0, 0.99, C.byref(d), C.byref(d), C.byref(d), C.byref(n)) == -1
    assert lib.gh_rd_steps(None, 1, 0.1, None) == -1
    assert lib.gh_phantom(4, 1, None, None, None, None) == -1


@pytest.mark.parametrize("real", ["f32", "f64"])
@pytest.mark.parametrize("args", [(0.013, 17.6, 300.0, 64, 0.9), (0.002, 17.6, 400.0, 128, 0.9),
                                  (0.015, 17.6, 400.0, 128, 0.99), (0.0013, 25.6, 300.0, 128, 0.99)])
def test_schedule_matches_oracle(oracle, real, args):
    assert gh.schedule(real, *args) == oracle.schedule(real, *args)


def test_phantom_is_seeded_and_physical():
    gm, wm, csf, sv = gh.phantom(48, 3)
    gm2, wm2, csf2, sv2 = gh.phantom(48, 3)
    assert np.array_equal(gm, gm2) and np.array_equal(wm, wm2) and sv == sv2
    gm3, _, _, _ = gh.phantom(48, 4)
    assert not np.array_equal(gm, gm3)
    tot = gm + wm + csf
    assert gm.min() >= 0 and wm.min() >= 0 and csf.min() >= 0 and tot.max() <= 1 + 1e-12
    i, j, k = sv
    assert wm[k, j, i] > 0.9
    assert tot[0, 0, 0] < 1e-6                       # corner: outside the head
    assert csf[24, 24, 2] > 0.4                      # just inside the head: outer CSF shell
    assert (wm > 0.5).sum() > 0 and (gm > 0.5).sum() > 0 and (csf > 0.5).sum() > 0


@pytest.mark.parametrize("real", ["f32", "f64"])
def test_anatomy_majority_vote_matches_oracle_rule(oracle, real):
    """The oracle's _ic restatement: CSF-dominant voxels become pure CSF, else
    p_w, p_g are renormalised to sum 1 (Glioma_ReactionDiffusion.cpp:143-152)."""
    gm, wm, csf, _ = gh.phantom(32, 9)
    pw, pg, pc = oracle.anatomy(gm, wm, csf, real)
    tissue = wm + gm
    csf_dom = csf > tissue
    assert np.all(pc[csf_dom] == 1) and np.all(pw[csf_dom] == 0)
    ok = (~csf_dom) & (tissue > 0)
    assert np.allclose((pw + pg)[ok], 1, atol=1e-6)


def test_gaussian_ic_gated_by_tissue():
    n = 16
    pw = np.zeros((n, n, n), np.float32)
    pg = np.zeros_like(pw)
    pw[4:12, 4:12, 4:12] = 1
    phi = gh.gaussian_ic("f32", pw, pg, (8, 8, 8), 1.0, 0.1)
    assert phi[8, 8, 8] == np.float32(0.1)
    assert phi[8, 8, 9] == np.float32(0.1 * np.exp(-0.5))
    assert np.all(phi[pw + pg == 0] == 0)


@pytest.mark.parametrize("real", ["f32", "f64"])
def test_rd_thresholds_equal_brute_force_predicates(real):
    """tau/sigma reproduce the reference predicates exactly: checked at tau,
    one ulp below it, and on random phi, with numpy's IEEE arithmetic."""
    t = np.float32 if real == "f32" else np.float64
    rng = np.random.default_rng(4)
    m = 4000
    pw = rng.uniform(-0.05, 1, m).astype(t)
    pg = rng.uniform(-0.05, 1, m).astype(t)
    pw[:500] = rng.uniform(0, 0.1, 500).astype(t)
    pg[:500] = rng.uniform(0, 0.01, 500).astype(t)
    pw[500:600] = 0
    pg[500:600] = 0
    tau, sig = gh.rd_thresholds(real, pw, pg)

    def outside(phi):
        return ((phi + pw).astype(t) + pg).astype(t) < t(0.1)

    def domain(phi):
        return ((pw + pg).astype(t) + phi).astype(t) > 0

    fin = np.isfinite(tau)
    assert fin[:600].all()
    at = tau.copy()
    below = np.nextafter(tau, t(-np.inf))
    assert not outside(at)[fin].any()
    assert outside(below)[fin].all()
    for _ in range(20):
        phi = rng.uniform(-0.2, 1.2, m).astype(t)
        assert np.array_equal(outside(phi), phi < tau)
        assert np.array_equal(domain(phi), phi > sig)
    assert np.array_equal(domain(np.zeros(m, t)), np.zeros(m, t) > sig)


def test_wrapper_shape_checks_refuse_before_the_c_call():
    """The C side reads n^3 values from every dense array: the wrappers refuse
    arrays of another shape (or a mismatched set) with ValueError first."""
    a = np.zeros((8, 8, 8))
    assert gh._cubes([a, a, a], "t") == 8
    assert gh._cubes([a], "t", 8) == 8
    for bad in ([a, np.zeros((8, 8, 4))], [np.zeros((8, 8, 8, 1))], [np.zeros((8, 8))]):
        with pytest.raises(ValueError):
            gh._cubes(bad, "t")
    with pytest.raises(ValueError):
        gh._cubes([a], "t", 16)
    r, n = gh._roi_rows([[1, 2, 3], [4, 5, 6]], "t")
    assert n == 2 and r.shape == (3, 2) and r[1, 0] == 2 and r.flags["C_CONTIGUOUS"]
    with pytest.raises(ValueError):
        gh._roi_rows([[1, 2], [3, 4]], "t")
    with pytest.raises(ValueError):
        gh.gaussian_ic("f32", np.zeros((8, 8, 8)), np.zeros((8, 8, 4)), (1, 1, 1))
