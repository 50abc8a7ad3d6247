"""Writes the golden fixtures under tests/golden/ from the reference-built
executables in oracle/_ref/ (built by oracle/Makefile from the reference
sources: Glioma/Operators/ReactionDiffusionOperator.h and
tools/pi4u_lite/Inference/Likelihood/source/*.cpp).

Run in the build container (where the reference tree exists):
    python tests/golden/make_golden.py
Each fixture holds inputs and the reference's outputs only.
"""
from __future__ import annotations

import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from conftest import REF_DIR, Oracle, run_rd_ref  # noqa: E402

import gliomasolver_amd as gh  # noqa: E402  (host-only phantom / schedule)


def random_anatomy(n, seed, t):
    """Anatomy that exercises every branch of the operator: pure CSF cells,
    tiny fractions (harmonic eps branch, chf near 0.1), negative fractions
    (in-domain test), ordinary tissue."""
    rng = np.random.default_rng(seed)
    pw = rng.uniform(0, 1, (n, n, n))
    pg = rng.uniform(0, 1, (n, n, n)) * (1 - pw)
    u = rng.uniform(0, 1, (n, n, n))
    pw[u < 0.12] = 0.0
    pg[u < 0.12] = 0.0
    tiny = (u >= 0.12) & (u < 0.2)
    pw[tiny] = rng.uniform(0, 0.06, tiny.sum())
    pg[tiny] = rng.uniform(0, 0.06, tiny.sum())
    eps = (u >= 0.2) & (u < 0.23)
    pw[eps] = 1e-9
    pg[eps] = 0.0
    neg = (u >= 0.23) & (u < 0.25)
    pw[neg] = -rng.uniform(0, 0.02, neg.sum())
    pg[neg] = rng.uniform(0, 0.01, neg.sum())
    phi = rng.uniform(0, 1, (n, n, n))
    phi[rng.uniform(0, 1, (n, n, n)) < 0.3] = 0.0
    return pw.astype(t), pg.astype(t), phi.astype(t)


def rd_case(name, real, pw, pg, phi, Dw, Dg, rho, h, dt, nsteps, tmp):
    t = np.float32 if real == "f32" else np.float64
    out = run_rd_ref(real, phi, pw, pg, Dw, Dg, rho, h, dt, nsteps, tmp)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), p_w=pw.astype(t), p_g=pg.astype(t), phi0=phi.astype(t),
                        phi_ref=out, Dw=Dw, Dg=Dg, rho=rho, h=h, dt=dt, nsteps=nsteps)
    print(name, "steps", nsteps, "max phi", float(out.max()), "sum", float(out.astype(np.float64).sum()))


def likelihood_ref(model, pet, t1c, flair, roi, stepPET, stepMRI, tmp):
    o = Oracle()
    for nm, a in (("HGG_data.dat", model), ("tumPET.dat", pet), ("tumT1c.dat", t1c), ("tumFLAIR.dat", flair)):
        nz, ny, nx = a.shape
        assert o.lib.gho_dat_write(os.path.join(tmp, nm).encode(), nx, ny, nz, o.p(np.ascontiguousarray(a))) == 0
    args = [os.path.join(REF_DIR, "likelihood"), "-stepPET", str(stepPET), "-stepMRI", str(stepMRI)]
    if roi is not None:
        # Matrix::D2D dump: int32 {1234, 2, nx, ny}, int32 type id 0, data x-fastest (i + nx*j)
        with open(os.path.join(tmp, "ROI.dat"), "wb") as f:
            f.write(np.array([1234, 2, roi.shape[0], 3, 0], np.int32).tobytes())
            f.write(np.ascontiguousarray(roi.T, np.float64).tobytes())
        args += ["-bROI", "1"]
    with open(os.path.join(tmp, "LikelihoodInput.txt"), "w") as f:
        f.write("0.000361\n0.85\n0.7\n0.25\n2\n")  # tools/pi4u_lite/Inference/Likelihood/makefile/LikelihoodInput.txt
    r = subprocess.run(args, cwd=tmp, capture_output=True, text=True, check=True)
    m = re.search(r"L_Pet=(\S+), L_T1=(\S+), L_T2 = (\S+)", r.stdout)
    comps = [float(v) for v in m.groups()]
    total = float(open(os.path.join(tmp, "Likelihood.txt")).read().split()[0])
    return total, comps


def main():
    if not os.path.exists(os.path.join(REF_DIR, "rd_ref_f32")):
        raise SystemExit("oracle/_ref not built (make -C oracle ref)")
    o = Oracle()
    with tempfile.TemporaryDirectory() as tmp:
        # 1-2: phantom, Gaussian IC at a seeded WM voxel, real schedule (T = 300 d)
        n = 32
        gm, wm, csf, sv = gh.phantom(n, 7)
        for real in ("f32", "f64"):
            pw, pg, _ = o.anatomy(gm, wm, csf, real)
            phi = gh.gaussian_ic(real, pw, pg, sv, 1.0, 0.1)
            Dw, Dg, dt, ns = gh.schedule(real, 0.013, 17.6, 300.0, n, 0.99)
            rd_case("rd_%s_phantom32" % real, real, pw, pg, phi, Dw, Dg, 0.025, 1.0 / n, dt, ns, tmp)
        # 3-4: random anatomy hitting every branch, fast reaction to reach the clamp
        n = 16
        for real in ("f32", "f64"):
            t = np.float32 if real == "f32" else np.float64
            pw, pg, phi = random_anatomy(n, 11, t)
            Dw, Dg, dt, _ = gh.schedule(real, 0.013, 17.6, 300.0, n, 0.99)
            rd_case("rd_%s_random16" % real, real, pw, pg, phi, Dw, Dg, 0.05, 1.0 / n, dt, 25, tmp)

        # likelihood: model = smooth tumour, data = thresholded masks with flips + PET noise
        rng = np.random.default_rng(5)
        n = 24
        z, y, x = np.meshgrid(*(np.arange(n),) * 3, indexing="ij")
        r2 = (x - 11.3) ** 2 + (y - 12.1) ** 2 + (z - 10.7) ** 2
        model = np.exp(-r2 / 30.0)
        model[3, 4, 5] = 0.7    # u == uc exactly: alpha = 1/2
        model[6, 7, 8] = 0.25
        t1c = (model >= 0.7).astype(np.float64)
        flair = (model >= 0.25).astype(np.float64)
        flip = rng.uniform(0, 1, model.shape) < 0.05
        t1c[flip] = 1 - t1c[flip]
        flip = rng.uniform(0, 1, model.shape) < 0.05
        flair[flip] = 1 - flair[flip]
        pet = np.where(t1c > 0, model + rng.normal(0, 0.05, model.shape), 0.0)
        roi = np.argwhere(flair > 0)[:, ::-1].astype(np.float64)  # (ix, iy, iz) rows
        cases = {}
        for tag, roi_, sp, sm in (("full", None, 1, 1), ("stride", None, 3, 1), ("roiA", roi, 3, 1),
                                  ("roiB", roi, 2, 2)):
            total, comps = likelihood_ref(model, pet, t1c, flair, roi_, sp, sm, tmp)
            cases[tag] = np.array([total] + comps + [sp, sm])
            print("likelihood", tag, total, comps)
        np.savez_compressed(os.path.join(HERE, "likelihood24.npz"), model=model, pet=pet, t1c=t1c, flair=flair,
                            roi=roi, **cases)


if __name__ == "__main__":
    main()
