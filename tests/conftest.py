"""Shared fixtures.  The oracle (oracle/liboracle.so) is the checker; tests are
the only product-side code allowed to load it."""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_DIR = os.path.join(ROOT, "oracle", "_ref")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


_vp, _i, _d, _i64 = C.c_void_p, C.c_int, C.c_double, C.c_int64


def _load_oracle():
    if not os.path.exists(ORACLE_SO):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "oracle"], check=True)
    lib = C.CDLL(ORACLE_SO)
    sig = {
        "gho_anatomy_f32": (None, [_i64, _vp, _vp, _vp, _vp, _vp, _vp]),
        "gho_anatomy_f64": (None, [_i64, _vp, _vp, _vp, _vp, _vp, _vp]),
        "gho_ic_f32": (None, [_i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
        "gho_ic_f64": (None, [_i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
        "gho_schedule_f32": (None, [_d, _d, _d, _i, _d, _vp, _vp, _vp, _vp]),
        "gho_schedule_f64": (None, [_d, _d, _d, _i, _d, _vp, _vp, _vp, _vp]),
        "gho_rd_steps_f32": (None, [_i, _vp, _vp, _vp, _vp, C.c_float, C.c_float, C.c_float, C.c_float, _d, _i64]),
        "gho_rd_steps_f64": (None, [_i, _vp, _vp, _vp, _vp, _d, _d, _d, _d, _d, _i64]),
        "gho_likelihood": (_d, [_i, _i, _i, _vp, _vp, _vp, _vp, _vp, _i, _d, _d, _d, _d, _d, _i, _i, _vp, _vp]),
        "gho_dat_write": (_i, [C.c_char_p, _i, _i, _i, _vp]),
        "gho_dat_read": (_i, [C.c_char_p, _i, _i, _i, _vp]),
        "gho_wavelet_taps": (_i, [_i, _i, _vp, _vp]),
        "gho_tlm_f64": (_d, [_i, _vp, _vp, _vp, _d, _d, _d, _d, _d, _i64, _d, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i,
                             _vp]),
        "gho_dloss_du": (None, [_i, _vp, _vp, _vp, _vp, _vp, _i, _i, _vp]),
        "gho_amr_lab_f32": (None, [_i, _vp, _i, _vp, _i, _vp]),
        "gho_amr_lab_f64": (None, [_i, _vp, _i, _vp, _i, _vp]),
        "gho_amr_fwt_f32": (_d, [_i, _vp]),
        "gho_amr_fwt_f64": (_d, [_i, _vp]),
        "gho_amr_split_f32": (None, [_i, _vp, _i, _vp]),
        "gho_amr_split_f64": (None, [_i, _vp, _i, _vp]),
        "gho_amr_rd_steps_f32": (None, [_i, _vp, _i, _vp, _vp, _vp, C.c_float, C.c_float, C.c_float, _d, _i64]),
        "gho_amr_rd_steps_f64": (None, [_i, _vp, _i, _vp, _vp, _vp, _d, _d, _d, _d, _i64]),
        "gho_amr_resample": (None, [_i, _vp, _vp, _i, _i, _vp]),
        "gho_amr_details_f32": (None, [_i, _vp, _i, _vp, _i, _vp, _vp]),
        "gho_amr_details_f64": (None, [_i, _vp, _i, _vp, _i, _vp, _vp]),
        "gho_amr_block_terms": (_i, [_i, _vp, _i, _i, _i, _i, _vp, _vp, _vp]),
        "gho_mrag_lab_f32": (_i, [_i, _vp, _i, _vp, _i, _vp]),
        "gho_mrag_lab_f64": (_i, [_i, _vp, _i, _vp, _i, _vp]),
        "gho_mrag_block_terms": (_i, [_i, _vp, _i, _i, _i, _i, _vp, _vp, _vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


class Oracle:
    """Thin numpy façade over liboracle.so."""

    def __init__(self):
        self.lib = _load_oracle()

    @staticmethod
    def p(a):
        return a.ctypes.data_as(C.c_void_p)

    def anatomy(self, gm, wm, csf, real):
        t = np.float32 if real == "f32" else np.float64
        pw, pg, pc = (np.zeros(gm.shape, t) for _ in range(3))
        f = self.lib.gho_anatomy_f32 if real == "f32" else self.lib.gho_anatomy_f64
        f(gm.size, self.p(gm), self.p(wm), self.p(csf), self.p(pw), self.p(pg), self.p(pc))
        return pw, pg, pc

    def ic(self, bpd, gm, wm, csf, tumor_ic, real):
        """_ic on a bpd^3-block grid from (nz, ny, nx) anatomy; returns pw, pg, pcsf, phi (n^3)."""
        t = np.float32 if real == "f32" else np.float64
        n = 8 * bpd
        out = [np.zeros((n, n, n), t) for _ in range(4)]
        gm, wm, csf = (np.ascontiguousarray(a, np.float64) for a in (gm, wm, csf))
        nz, ny, nx = gm.shape
        ic = None if tumor_ic is None else (C.c_double * 3)(*tumor_ic)
        f = self.lib.gho_ic_f32 if real == "f32" else self.lib.gho_ic_f64
        f(bpd, nx, ny, nz, self.p(gm), self.p(wm), self.p(csf), ic, *[self.p(a) for a in out])
        return tuple(out)

    def schedule(self, real, Dw_phys, L, tend, gpd, cfl):
        t = C.c_float if real == "f32" else C.c_double
        Dw, Dg, dt, ns = t(), t(), t(), C.c_int64()
        f = self.lib.gho_schedule_f32 if real == "f32" else self.lib.gho_schedule_f64
        f(Dw_phys, L, tend, gpd, cfl, C.byref(Dw), C.byref(Dg), C.byref(dt), C.byref(ns))
        return Dw.value, Dg.value, dt.value, ns.value

    def rd_steps(self, phi, pw, pg, Dw, Dg, rho, h, dt, nsteps):
        """Returns phi after nsteps (input not modified)."""
        real = "f32" if phi.dtype == np.float32 else "f64"
        phi = phi.copy()
        dphi = np.zeros_like(phi)
        f = self.lib.gho_rd_steps_f32 if real == "f32" else self.lib.gho_rd_steps_f64
        f(phi.shape[0], self.p(phi), self.p(dphi), self.p(pw), self.p(pg), Dw, Dg, rho, h, dt, nsteps)
        return phi

    def likelihood(self, model, pet, t1c, flair, roi=None, PETsigma2=0.000361, PETscale=0.85, T1uc=0.7, T2uc=0.25,
                   slope=2.0, stepPET=1, stepTi=1):
        nx, ny, nz = model.shape[2], model.shape[1], model.shape[0]
        out = np.zeros(3)
        npet = C.c_int64()
        r = None if roi is None else np.ascontiguousarray(roi.T, np.float64)  # column-major (i + n*col)
        total = self.lib.gho_likelihood(nx, ny, nz, self.p(model), self.p(pet), self.p(t1c), self.p(flair),
                                        None if r is None else self.p(r), 0 if roi is None else roi.shape[0],
                                        PETsigma2, PETscale, T1uc, T2uc, slope, stepPET, stepTi, self.p(out),
                                        C.byref(npet))
        return total, out, npet.value

    LIK = (0.000361, 0.85, 0.7, 0.25, 2.0)

    def tlm(self, phi0, pw, pg, Dw, Dg, rho, h, T, N, sigma, x0, direction, t1c, flair, pet, lik=LIK, stepPET=1,
            stepTi=1):
        """(J, dJ along direction) of the FP64 tangent-linear model."""
        n = phi0.shape[0]
        x0a = np.asarray(x0, np.float64)
        da = np.asarray(direction, np.float64)
        la = np.asarray(lik, np.float64)
        loss = C.c_double()
        dj = self.lib.gho_tlm_f64(n, self.p(phi0), self.p(pw), self.p(pg), Dw, Dg, rho, h, T, N, sigma, self.p(x0a),
                                  self.p(da), self.p(t1c), self.p(flair), self.p(pet), self.p(la), stepPET, stepTi,
                                  C.byref(loss))
        return loss.value, dj

    def dloss_du(self, u, t1c, flair, pet, lik=LIK, stepPET=1, stepTi=1):
        g = np.zeros_like(u)
        la = np.asarray(lik, np.float64)
        self.lib.gho_dloss_du(u.shape[0], self.p(u), self.p(t1c), self.p(flair), self.p(pet), self.p(la), stepPET,
                              stepTi, self.p(g))
        return g


    # ---- wavelet-adaptive grid: leaves = canonical rows (level, ix, iy, iz) ----
    @staticmethod
    def _rows(leaves):
        return np.ascontiguousarray(leaves, np.int32).reshape(-1, 4)

    def amr_lab(self, leaves, order, field, leaf):
        r = self._rows(leaves)
        f = np.ascontiguousarray(field)
        n = 13 if order == 4 else 10
        lab = np.zeros(n ** 3, f.dtype)
        fn = self.lib.gho_amr_lab_f32 if f.dtype == np.float32 else self.lib.gho_amr_lab_f64
        fn(len(r), self.p(r), order, self.p(f), leaf, self.p(lab))
        return lab.reshape(n, n, n)

    def mrag_lab(self, leaves, order, field, leaf):
        """The lab as MRAG's instruction lists give it (gh_oracle_mrag.c)."""
        r = self._rows(leaves)
        f = np.ascontiguousarray(field)
        n = 13 if order == 4 else 10
        lab = np.zeros(n ** 3, f.dtype)
        fn = self.lib.gho_mrag_lab_f32 if f.dtype == np.float32 else self.lib.gho_mrag_lab_f64
        rc = fn(len(r), self.p(r), order, self.p(f), leaf, self.p(lab))
        assert rc == 0, "MRAG restatement refused the grid around leaf %d" % leaf
        return lab.reshape(n, n, n)

    def block_terms(self, which, leaves, order, leaf, f32=True, cap=400000):
        """Per lab position of `leaf` (N^3, x fastest): counts, then the
        sources and Real weights of all lists in order; which = 'amr' (the
        recursion) or 'mrag' (MRAG's instruction lists)."""
        r = self._rows(leaves)
        n = 13 if order == 4 else 10
        counts = np.zeros(n ** 3, np.int32)
        src = np.zeros(cap, np.int32)
        w = np.zeros(cap, np.float64)
        fn = self.lib.gho_amr_block_terms if which == "amr" else self.lib.gho_mrag_block_terms
        tot = fn(len(r), self.p(r), order, leaf, int(f32), cap, self.p(counts), self.p(src), self.p(w))
        assert 0 <= tot <= cap, tot
        return counts, src[:tot].copy(), w[:tot].copy()

    def amr_details(self, leaves, order, field, subset):
        """FWT details of the listed leaves (lab + BlockFWT report each)."""
        r = self._rows(leaves)
        f = np.ascontiguousarray(field)
        lst = np.ascontiguousarray(subset, np.int32)
        out = np.zeros(len(lst), np.float64)
        fn = self.lib.gho_amr_details_f32 if f.dtype == np.float32 else self.lib.gho_amr_details_f64
        fn(len(r), self.p(r), order, self.p(f), len(lst), self.p(lst), self.p(out))
        return out

    def amr_fwt(self, order, lab):
        lab = np.ascontiguousarray(lab)
        fn = self.lib.gho_amr_fwt_f32 if lab.dtype == np.float32 else self.lib.gho_amr_fwt_f64
        return fn(order, self.p(lab))

    def amr_split(self, order, lab, octant):
        lab = np.ascontiguousarray(lab)
        out = np.zeros(512, lab.dtype)
        fn = self.lib.gho_amr_split_f32 if lab.dtype == np.float32 else self.lib.gho_amr_split_f64
        fn(order, self.p(lab), octant, self.p(out))
        return out.reshape(8, 8, 8)

    def amr_rd_steps(self, leaves, order, phi, pw, pg, Dw, Dg, rho, dt, nsteps):
        r = self._rows(leaves)
        phi = np.ascontiguousarray(phi).copy()
        fn = self.lib.gho_amr_rd_steps_f32 if phi.dtype == np.float32 else self.lib.gho_amr_rd_steps_f64
        fn(len(r), self.p(r), order, self.p(phi), self.p(np.ascontiguousarray(pw)), self.p(np.ascontiguousarray(pg)),
           Dw, Dg, rho, dt, nsteps)
        return phi

    def amr_resample(self, leaves, phi, gpd):
        r = self._rows(leaves)
        out = np.zeros((gpd, gpd, gpd))
        ph = np.ascontiguousarray(phi, np.float64)
        self.lib.gho_amr_resample(len(r), self.p(r), self.p(ph), int(phi.dtype == np.float32), gpd, self.p(out))
        return out


@pytest.fixture(scope="session")
def oracle():
    return Oracle()


def ref_available(name: str) -> bool:
    return os.path.exists(os.path.join(REF_DIR, name))


def run_rd_ref(real: str, phi, pw, pg, Dw, Dg, rho, h, dt, nsteps, workdir):
    """Run oracle/_ref/rd_ref_{f32,f64} (the reference operator) on dense arrays."""
    exe = os.path.join(REF_DIR, "rd_ref_" + real)
    t = np.float32 if real == "f32" else np.float64
    n = phi.shape[0]
    inp = os.path.join(workdir, "rd_in.bin")
    out = os.path.join(workdir, "rd_out.bin")
    with open(inp, "wb") as f:
        f.write(np.array([n], np.int32).tobytes())
        f.write(np.array([nsteps], np.int64).tobytes())
        f.write(np.array([Dw, Dg, rho, h, dt], np.float64).tobytes())
        for a in (pw, pg, phi):
            f.write(np.ascontiguousarray(a, t).tobytes())
    subprocess.run([exe, inp, out], check=True, capture_output=True)
    return np.fromfile(out, t).reshape(phi.shape)


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
