"""gliomasolver_amd -- MI355X-native reaction-diffusion hot path of GliomaSolver.

Python mirror of the C ABI in include/gliomahip.h.  All compute runs in
libgliomahip.so (HIP, gfx950); this module only moves numpy arrays across the
ABI.  Names follow the reference (Glioma/Glioma_ReactionDiffusion.cpp,
Glioma/Operators/ReactionDiffusionOperator.h).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi
from ._capi import (FIELD_DPHIDT, FIELD_PCSF, FIELD_PG, FIELD_PHI, FIELD_PW, GH_REAL_F32, GH_REAL_F64,
                    GHError, check, load)

__all__ = ["Context", "GHError", "phantom", "schedule", "gaussian_ic", "rd_thresholds", "real_kind", "np_real", "load",
           "FIELD_PHI", "FIELD_PW", "FIELD_PG", "FIELD_PCSF", "FIELD_DPHIDT", "GH_REAL_F32", "GH_REAL_F64"]


def real_kind(real) -> int:
    if real in ("f32", "float", "float32", np.float32, GH_REAL_F32):
        return GH_REAL_F32
    if real in ("f64", "double", "float64", np.float64, GH_REAL_F64):
        return GH_REAL_F64
    raise ValueError("real must be f32 or f64, got %r" % (real,))


def np_real(kind: int):
    return np.float32 if kind == GH_REAL_F32 else np.float64


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _cubes(arrs, what: str, n: int = None) -> int:
    """Side n of arrays that must all be (n, n, n); the C side reads n^3 values
    from each, so a wrong shape is refused here rather than read past."""
    if n is None:
        n = arrs[0].shape[0] if arrs[0].ndim == 3 else -1
    for a in arrs:
        if a.shape != (n, n, n):
            raise ValueError("%s: every array must have shape (%d, %d, %d), got %s" % (what, n, n, n, a.shape))
    return n


def _roi_rows(roi, what: str):
    r = np.asarray(roi, np.float64)
    if r.ndim != 2 or r.shape[1] != 3:
        raise ValueError("%s: roi must have shape (N, 3), got %s" % (what, r.shape))
    return np.ascontiguousarray(r.T), r.shape[0]


def phantom(n: int, seed: int):
    """Seeded synthetic brain: returns (gm, wm, csf, seed_voxel) as n^3 float64 arrays (x fastest)."""
    lib = load()
    gm = np.zeros((n, n, n), np.float64)
    wm = np.zeros_like(gm)
    csf = np.zeros_like(gm)
    sv = (C.c_int * 3)()
    check(lib.gh_phantom(n, C.c_uint64(seed), _ptr(gm), _ptr(wm), _ptr(csf), sv))
    return gm, wm, csf, (sv[0], sv[1], sv[2])


def schedule(real, Dw_phys: float, L: float, tend: float, gpd: int, cfl: float = 0.99):
    """Rescaled Dw, Dg, dt and step count of Glioma_ReactionDiffusion::run (all in Real)."""
    lib = load()
    Dw, Dg, dt = C.c_double(), C.c_double(), C.c_double()
    ns = C.c_int64()
    check(lib.gh_schedule(real_kind(real), Dw_phys, L, tend, gpd, cfl, C.byref(Dw), C.byref(Dg), C.byref(dt),
                          C.byref(ns)))
    return Dw.value, Dg.value, dt.value, ns.value


def gaussian_ic(real, p_w: np.ndarray, p_g: np.ndarray, center, sigma: float = 1.0, peak: float = 0.1):
    k = real_kind(real)
    t = np_real(k)
    pw = np.ascontiguousarray(p_w, t)
    pg = np.ascontiguousarray(p_g, t)
    n = _cubes([pw, pg], "gaussian_ic")
    out = np.zeros_like(pw)
    c = (C.c_double * 3)(*[float(v) for v in center])
    check(load().gh_gaussian_ic(k, n, _ptr(pw), _ptr(pg), c, sigma, peak, _ptr(out)))
    return out


def rd_thresholds(real, p_w: np.ndarray, p_g: np.ndarray):
    """(tau, sigma) per cell: neighbour-outside iff phi < tau, in-domain iff phi > sigma."""
    t = np_real(real_kind(real))
    pw = np.ascontiguousarray(p_w, t)
    pg = np.ascontiguousarray(p_g, t)
    tau = np.zeros_like(pw)
    sig = np.zeros_like(pw)
    check(load().gh_rd_thresholds(real_kind(real), pw.size, _ptr(pw), _ptr(pg), _ptr(tau), _ptr(sig)))
    return tau, sig


def canonical_leaves(rows) -> np.ndarray:
    """Leaf rows (level, ix, iy, iz) sorted into the canonical (level, iz, iy, ix) order."""
    r = np.asarray(rows, np.int32).reshape(-1, 4)
    return np.ascontiguousarray(r[np.lexsort((r[:, 1], r[:, 2], r[:, 3], r[:, 0]))])


def plan_neighbours(rows, leaf: int):
    r = canonical_leaves(rows)
    out = np.zeros((128, 2), np.int32)
    n = C.c_int()
    check(load().gh_plan_neighbours(len(r), _ptr(r), leaf, C.byref(n), _ptr(out)))
    return [(int(a), bool(b)) for a, b in out[:n.value]]


def plan_refine(rows, flagged, max_level: int):
    r = canonical_leaves(rows)
    f = np.ascontiguousarray(flagged, np.uint8)
    out = np.zeros(len(r), np.int32)
    n = C.c_int()
    check(load().gh_plan_refine(len(r), _ptr(r), _ptr(f), max_level, C.byref(n), _ptr(out)))
    return sorted(int(v) for v in out[:n.value])


def plan_compress(rows, flagged, min_level: int = 0):
    r = canonical_leaves(rows)
    f = np.ascontiguousarray(flagged, np.uint8)
    out = np.zeros((len(r) // 8 + 1, 4), np.int32)
    n = C.c_int()
    check(load().gh_plan_compress(len(r), _ptr(r), _ptr(f), min_level, C.byref(n), _ptr(out)))
    return sorted(tuple(int(v) for v in row) for row in out[:n.value])


class Context:
    """One device context (gh_ctx): grid topology plus block-major fields in HBM."""

    def __init__(self, device: int = 0, real="f32"):
        self.lib = load()
        self.kind = real_kind(real)
        self.dtype = np_real(self.kind)
        h = C.c_void_p()
        check(self.lib.gh_ctx_create(device, self.kind, C.byref(h)))
        self.h = h
        self.n = 0

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.gh_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- grid ----
    def grid_init_uniform(self, bpd: int):
        check(self.lib.gh_grid_init_uniform(self.h, bpd))
        self.n = 8 * bpd

    def nblocks(self) -> int:
        v = C.c_int()
        check(self.lib.gh_grid_nblocks(self.h, C.byref(v)))
        return v.value

    def blocks(self):
        nb = self.nblocks()
        lev = np.zeros(nb, np.int32)
        idx = np.zeros((nb, 3), np.int32)
        check(self.lib.gh_grid_blocks(self.h, _ptr(lev), _ptr(idx)))
        return lev, idx

    # ---- wavelet-adaptive grid ----
    def grid_configure(self, wavelet_order: int = 4, max_level: int = 4, min_level: int = 0):
        check(self.lib.gh_grid_configure(self.h, wavelet_order, max_level, min_level))

    def grid_init_leaves(self, levels, indices):
        lev = np.ascontiguousarray(levels, np.int32)
        idx = np.ascontiguousarray(indices, np.int32).reshape(-1, 3)
        check(self.lib.gh_grid_init_leaves(self.h, len(lev), _ptr(lev), _ptr(idx)))
        self.n = 0

    def fwt_details(self) -> np.ndarray:
        out = np.zeros(self.nblocks(), np.float64)
        check(self.lib.gh_fwt_details(self.h, _ptr(out)))
        return out

    def refine(self, tol: float, max_loops: int = 1) -> int:
        n = C.c_int()
        check(self.lib.gh_refine(self.h, tol, max_loops, C.byref(n)))
        return n.value

    def compress(self, tol: float, max_loops: int = -1) -> int:
        n = C.c_int()
        check(self.lib.gh_compress(self.h, tol, max_loops, C.byref(n)))
        return n.value

    def refine_blocks(self, block_ids) -> int:
        """Grid::refine(set<int>): split the listed blocks with the refiner's
        plan (space extension, maxLevel, 2:1 chain reaction)."""
        ids = np.ascontiguousarray(np.unique(np.asarray(block_ids, np.int64)), np.int32)
        n = C.c_int()
        check(self.lib.gh_grid_refine_blocks(self.h, len(ids), _ptr(ids), C.byref(n)))
        return n.value

    def compress_blocks(self, block_ids) -> int:
        """Grid::compress(set<int>, int&): collapse every parent whose eight
        children are listed and pass the compressor's level-jump test."""
        ids = np.ascontiguousarray(np.unique(np.asarray(block_ids, np.int64)), np.int32)
        n = C.c_int()
        check(self.lib.gh_grid_compress_blocks(self.h, len(ids), _ptr(ids), C.byref(n)))
        return n.value

    ADAPT_PHASES = ("ref_fwt", "ref", "comp_fwt", "comp", "plan", "move", "topology", "statics")

    def adapt_profile(self, reset: bool = False) -> dict:
        """Host wall time (ms) of gh_refine / gh_compress per phase since the
        last reset (MRAG Profiler agents, then plan/move/topology/statics)."""
        ms = np.zeros(8)
        check(self.lib.gh_adapt_profile(self.h, _ptr(ms), int(reset)))
        return dict(zip(self.ADAPT_PHASES, ms.tolist()))

    def adapt_counters(self, reset: bool = False) -> dict:
        """How ghosts were built since the last reset: lab batches on the
        device / from host tables, face tables on the device / on the host."""
        n = np.zeros(4, np.int64)
        check(self.lib.gh_adapt_counters(self.h, _ptr(n), int(reset)))
        return dict(zip(("labs_device", "labs_host", "faces_device", "faces_host"), n.tolist()))

    def set_field_blocks(self, field: int, values: np.ndarray):
        v = np.ascontiguousarray(values, self.dtype).reshape(self.nblocks(), 8, 8, 8)
        check(self.lib.gh_set_field_blocks(self.h, field, _ptr(v)))

    def get_field_blocks(self, field: int) -> np.ndarray:
        out = np.zeros((self.nblocks(), 8, 8, 8), self.dtype)
        check(self.lib.gh_get_field_blocks(self.h, field, _ptr(out)))
        return out

    def uq_resample(self, gpd: int) -> np.ndarray:
        out = np.zeros((gpd, gpd, gpd), np.float64)
        check(self.lib.gh_uq_resample(self.h, gpd, _ptr(out)))
        return out

    # ---- fields ----
    def set_anatomy(self, gm, wm, csf):
        arrs = [np.ascontiguousarray(a, np.float64) for a in (gm, wm, csf)]
        n = _cubes(arrs, "set_anatomy", self.n)
        check(self.lib.gh_set_anatomy(self.h, n, *[_ptr(a) for a in arrs]))

    def ic_from_anatomy(self, gm, wm, csf, tumor_ic=None):
        """Glioma_ReactionDiffusion::_ic for anatomy arrays of shape (nz, ny, nx)."""
        arrs = [np.ascontiguousarray(a, np.float64) for a in (gm, wm, csf)]
        if arrs[0].ndim != 3 or any(a.shape != arrs[0].shape for a in arrs):
            raise ValueError("ic_from_anatomy: gm, wm, csf must be 3-D arrays of one shape")
        nz, ny, nx = arrs[0].shape
        ic = None if tumor_ic is None else (C.c_double * 3)(*[float(v) for v in tumor_ic])
        check(self.lib.gh_ic_from_anatomy(self.h, nx, ny, nz, *[_ptr(a) for a in arrs], ic))

    def set_field(self, field: int, values: np.ndarray):
        v = np.ascontiguousarray(values, self.dtype)
        check(self.lib.gh_set_field_dense(self.h, field, _cubes([v], "set_field", self.n), _ptr(v)))

    def get_field(self, field: int) -> np.ndarray:
        out = np.zeros((self.n, self.n, self.n), self.dtype)
        check(self.lib.gh_get_field_dense(self.h, field, self.n, _ptr(out)))
        return out

    # ---- reaction-diffusion ----
    def rd_prepare(self, Dw: float, Dg: float, rho: float):
        check(self.lib.gh_rd_prepare(self.h, Dw, Dg, rho))

    def rd_rhs(self):
        check(self.lib.gh_rd_rhs(self.h))

    def rd_update(self, dt: float):
        check(self.lib.gh_rd_update(self.h, dt))

    def rd_steps(self, nsteps: int, dt: float, timed: bool = False):
        ms = C.c_double()
        check(self.lib.gh_rd_steps(self.h, nsteps, dt, C.byref(ms) if timed else None))
        return ms.value if timed else None

    def rd_set_tile(self, blocks_per_workgroup: int):
        check(self.lib.gh_rd_set_tile(self.h, blocks_per_workgroup))

    def rd_autotune(self, dt: float, trial_steps: int = 50) -> int:
        best = C.c_int()
        check(self.lib.gh_rd_autotune(self.h, dt, trial_steps, C.byref(best)))
        return best.value

    def rd_schedule_info(self):
        """(tile, clamped_pairs): the step schedule and whether pairs of steps
        run the clamped-state kernel."""
        tile, cp = C.c_int(), C.c_int()
        check(self.lib.gh_rd_schedule_info(self.h, C.byref(tile), C.byref(cp)))
        return tile.value, bool(cp.value)

    def rd_zero_flux_blocks(self) -> int:
        """Blocks without diffusion, advanced by the reaction update alone."""
        n = C.c_int()
        check(self.lib.gh_rd_zero_flux_blocks(self.h, C.byref(n)))
        return n.value

    # ---- likelihood (HGG_Likelihood) ----
    LIK_DEFAULT = (0.000361, 0.85, 0.7, 0.25, 2.0)  # LikelihoodInput.txt

    def lik_set_data(self, t1c, flair, pet, roi=None):
        arrs = [np.ascontiguousarray(a, np.float64) for a in (t1c, flair, pet)]
        n = _cubes(arrs, "lik_set_data", self.n)
        r, nroi = (None, 0) if roi is None else _roi_rows(roi, "lik_set_data")
        check(self.lib.gh_lik_set_data(self.h, n, *[_ptr(a) for a in arrs], None if r is None else _ptr(r), nroi))

    def likelihood(self, params=LIK_DEFAULT, stepPET=1, stepTi=1, use_roi=False):
        """Returns (-logL, (L_PET, L_T1, L_T2), N_PET) for the current phi."""
        p = (C.c_double * 5)(*params)
        out = (C.c_double * 4)()
        npet = C.c_int64()
        check(self.lib.gh_likelihood(self.h, p, stepPET, stepTi, int(use_roi), out, C.byref(npet)))
        return out[0], (out[1], out[2], out[3]), npet.value

    def likelihood_dense(self, model, t1c, flair, pet, roi=None, params=LIK_DEFAULT, stepPET=1, stepTi=1):
        arrs = [np.ascontiguousarray(a, np.float64) for a in (model, t1c, flair, pet)]
        if arrs[0].ndim != 3 or any(a.shape != arrs[0].shape for a in arrs):
            raise ValueError("likelihood_dense: model, t1c, flair, pet must be 3-D arrays of one shape")
        nz, ny, nx = arrs[0].shape
        r, nroi = (None, 0) if roi is None else _roi_rows(roi, "likelihood_dense")
        p = (C.c_double * 5)(*params)
        out = (C.c_double * 4)()
        npet = C.c_int64()
        check(self.lib.gh_likelihood_dense(self.h, nx, ny, nz, *[_ptr(a) for a in arrs],
                                           None if r is None else _ptr(r), nroi, p,
                                           stepPET, stepTi, out, C.byref(npet)))
        return out[0], (out[1], out[2], out[3]), npet.value

    def rd_adjoint(self, nsteps: int, T: float, ic_center, ic_sigma: float = 1.0, ckpt_every: int = 64,
                   params=LIK_DEFAULT, stepPET=1, stepTi=1, use_roi=False):
        """Forward + adjoint from the current phi: returns (loss[4], grad[6])
        with grad = dJ/d(Dw, rho, T, x0, y0, z0), Dw as passed to rd_prepare."""
        p = (C.c_double * 5)(*params)
        ic = (C.c_double * 3)(*[float(v) for v in ic_center])
        loss = (C.c_double * 4)()
        grad = (C.c_double * 6)()
        check(self.lib.gh_rd_adjoint(self.h, nsteps, T, ckpt_every, p, stepPET, stepTi, int(use_roi), ic, ic_sigma,
                                     loss, grad))
        return np.array(loss[:]), np.array(grad[:])

    def checkpoint_save(self, slot: int):
        check(self.lib.gh_checkpoint_save(self.h, slot))

    def checkpoint_load(self, slot: int):
        check(self.lib.gh_checkpoint_load(self.h, slot))

    def sync(self):
        check(self.lib.gh_sync(self.h))


class Batch:
    """A population of independent uniform solves (gh_batch): one device solve
    of many samples, each bit-identical to solving it alone with Context."""

    def __init__(self, device: int = 0, real="f32"):
        self.lib = load()
        self.kind = real_kind(real)
        self.dtype = np_real(self.kind)
        h = C.c_void_p()
        check(self.lib.gh_batch_create(device, self.kind, C.byref(h)))
        self.h = h
        self.sizes = []
        self.ns = 0
        self.anat_of = []

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.gh_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def add_anatomy(self, bpd: int, gm, wm, csf) -> int:
        """GM/WM/CSF of shape (nz, ny, nx), any size; grid of bpd^3 blocks."""
        arrs = [np.ascontiguousarray(a, np.float64) for a in (gm, wm, csf)]
        if arrs[0].ndim != 3 or any(a.shape != arrs[0].shape for a in arrs):
            raise ValueError("add_anatomy: gm, wm, csf must be 3-D arrays of one shape")
        nz, ny, nx = arrs[0].shape
        aid = C.c_int()
        check(self.lib.gh_batch_add_anatomy(self.h, bpd, nx, ny, nz, *[_ptr(a) for a in arrs], C.byref(aid)))
        self.sizes.append(8 * bpd)
        return aid.value

    def set_observations(self, anatomy: int, t1c, flair, pet, roi=None):
        if not 0 <= anatomy < len(self.sizes):
            raise ValueError("set_observations: unknown anatomy %d" % anatomy)
        arrs = [np.ascontiguousarray(a, np.float64) for a in (t1c, flair, pet)]
        _cubes(arrs, "set_observations", self.sizes[anatomy])
        r, nroi = (None, 0) if roi is None else _roi_rows(roi, "set_observations")
        check(self.lib.gh_batch_set_observations(self.h, anatomy, *[_ptr(a) for a in arrs],
                                                 None if r is None else _ptr(r), nroi))

    def set_samples(self, anatomy, params, cfl: float = 0.99):
        """params: (N, 6) rows {Dw, rho, Tend, icx, icy, icz}."""
        a = np.ascontiguousarray(anatomy, np.int32)
        p = np.ascontiguousarray(params, np.float64).reshape(len(a), 6)
        check(self.lib.gh_batch_set_samples(self.h, len(a), _ptr(a), _ptr(p), cfl))
        self.ns = len(a)
        self.anat_of = [int(v) for v in a]

    def run(self) -> float:
        ms = C.c_double()
        check(self.lib.gh_batch_run(self.h, C.byref(ms)))
        return ms.value

    def set_adjoint_samples(self, anatomy, params, nsteps, ic_sigma: float = 1.0, ic_peak: float = 0.1):
        """params: (N, 6) rows {Dw (/L^2), rho, T, x0, y0, z0 (cells)}; nsteps: N per sample."""
        a = np.ascontiguousarray(anatomy, np.int32)
        p = np.ascontiguousarray(params, np.float64).reshape(len(a), 6)
        n = np.ascontiguousarray(nsteps, np.int64)
        check(self.lib.gh_batch_set_adjoint_samples(self.h, len(a), _ptr(a), _ptr(p), _ptr(n), ic_sigma, ic_peak))
        self.ns = len(a)
        self.anat_of = [int(v) for v in a]

    def adjoint(self, lik_params, ckpt_every: int = 64, stepPET=1, stepTi=1, use_roi=False):
        """Forward + loss + adjoint of every sample: (loss (N, 4), grad (N, 6), device ms)."""
        lp = np.asarray(lik_params, np.float64)
        if lp.ndim == 1:
            lp = np.tile(lp, (self.ns, 1))
        lp = np.ascontiguousarray(lp)
        loss = np.zeros((self.ns, 4))
        grad = np.zeros((self.ns, 6))
        ms = C.c_double()
        check(self.lib.gh_batch_adjoint(self.h, ckpt_every, _ptr(lp), stepPET, stepTi, int(use_roi), _ptr(loss),
                                        _ptr(grad), C.byref(ms)))
        return loss, grad, ms.value

    def likelihood(self, lik_params, stepPET=1, stepTi=1, use_roi=False) -> np.ndarray:
        """(N, 4) rows {-logL, L_PET, L_T1, L_T2}; lik_params (N, 5) or one row for all."""
        lp = np.asarray(lik_params, np.float64)
        if lp.ndim == 1:
            lp = np.tile(lp, (self.ns, 1))
        lp = np.ascontiguousarray(lp)
        out = np.zeros((self.ns, 4))
        check(self.lib.gh_batch_likelihood(self.h, _ptr(lp), stepPET, stepTi, int(use_roi), _ptr(out)))
        return out

    def sample_info(self, i: int):
        Dw, Dg, dt, ns = C.c_double(), C.c_double(), C.c_double(), C.c_int64()
        check(self.lib.gh_batch_sample_info(self.h, i, C.byref(Dw), C.byref(Dg), C.byref(dt), C.byref(ns)))
        return Dw.value, Dg.value, dt.value, ns.value

    def get_phi(self, i: int) -> np.ndarray:
        n = self.sizes[self.anat_of[i]]
        out = np.zeros((n, n, n), self.dtype)
        check(self.lib.gh_batch_get_phi(self.h, i, _ptr(out)))
        return out

    def get_anatomy(self, anatomy: int):
        """(p_w, p_g) of an anatomy on its grid, (n, n, n) each."""
        if not 0 <= anatomy < len(self.sizes):
            raise ValueError("get_anatomy: unknown anatomy %d" % anatomy)
        n = self.sizes[anatomy]
        pw, pg = np.zeros((n, n, n), self.dtype), np.zeros((n, n, n), self.dtype)
        check(self.lib.gh_batch_get_anatomy(self.h, anatomy, _ptr(pw), _ptr(pg)))
        return pw, pg


class Tmcmc:
    """The TMCMC sampler (gh_tmcmc): ask for candidates, evaluate them (e.g.
    with Batch), submit their log-likelihoods, until finished."""

    def __init__(self, lo, hi, pop_size, max_stages=1000, min_chain=0, max_chain=1, tol_cov=1.0, bbeta=0.2,
                 seed=280675):
        self.lib = load()
        self.lo = np.ascontiguousarray(lo, np.float64)
        self.hi = np.ascontiguousarray(hi, np.float64)
        self.nth = len(self.lo)
        h = C.c_void_p()
        check(self.lib.gh_tmcmc_create(self.nth, _ptr(self.lo), _ptr(self.hi), pop_size, max_stages, min_chain,
                                       max_chain, tol_cov, bbeta, seed, C.byref(h)))
        self.h = h
        self.pop = pop_size

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.gh_tmcmc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def candidates(self) -> np.ndarray:
        n = C.c_int()
        check(self.lib.gh_tmcmc_candidates(self.h, C.byref(n), None))
        out = np.zeros((n.value, self.nth))
        check(self.lib.gh_tmcmc_candidates(self.h, C.byref(n), _ptr(out)))
        return out

    def submit(self, loglik):
        f = np.ascontiguousarray(loglik, np.float64)
        check(self.lib.gh_tmcmc_submit(self.h, len(f), _ptr(f)))

    def state(self):
        g, p, cov, ev, done = C.c_int(), C.c_double(), C.c_double(), C.c_double(), C.c_int()
        check(self.lib.gh_tmcmc_state(self.h, C.byref(g), C.byref(p), C.byref(cov), C.byref(ev), C.byref(done)))
        return dict(gen=g.value, p=p.value, cov=cov.value, log_evidence=ev.value, done=bool(done.value))

    def population(self):
        n = C.c_int()
        check(self.lib.gh_tmcmc_population(self.h, C.byref(n), None, None))
        x = np.zeros((n.value, self.nth))
        f = np.zeros(n.value)
        check(self.lib.gh_tmcmc_population(self.h, C.byref(n), _ptr(x), _ptr(f)))
        return x, f


def tmcmc_next_exponent(loglik, p: float, tol_cov: float = 1.0):
    f = np.ascontiguousarray(loglik, np.float64)
    pn, cov = C.c_double(), C.c_double()
    check(load().gh_tmcmc_next_exponent(len(f), _ptr(f), p, tol_cov, C.byref(pn), C.byref(cov)))
    return pn.value, cov.value


def glioma_params(x, real="f32"):
    """fitfun_glioma.c's map of 11 log-space parameters to the solver's
    {Dw, rho, Tend, icx, icy, icz} and the likelihood's 5 inputs, through the
    reference's text files.  x: (N, 11)."""
    xa = np.ascontiguousarray(np.atleast_2d(x), np.float64)
    s = np.zeros((len(xa), 6))
    lk = np.zeros((len(xa), 5))
    check(load().gh_glioma_params(len(xa), _ptr(xa), real_kind(real), _ptr(s), _ptr(lk)))
    return s, lk


def glioma_fitness(neg_loglik):
    """-logL -> the log-likelihood TMCMC sees (Likelihood.txt round trip, 1e12 penalty)."""
    v = np.ascontiguousarray(neg_loglik, np.float64)
    out = np.zeros_like(v)
    check(load().gh_glioma_fitness(len(v), _ptr(v), _ptr(out)))
    return out
