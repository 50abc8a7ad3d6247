"""GPU checks of the drop-in boundary for the reference's own drivers:

1. gh_ic_from_anatomy (Glioma_ReactionDiffusion::_ic) bit-exact against the
   oracle's restatement, on anatomy of a different, non-cubic size;
2. `brain -model RD -UQ 1` (reads TumorIC.txt / InputParameters.txt and
   {GM,WM,CSF}.dat, adapts at dump times, writes HGG_data.dat) bit-exact
   against the same run driven step by step from Python through the C ABI,
   and against `rd_pipeline`, the reference-structured per-step driver over
   the C++ mirror types;
3. `likelihood` (LikelihoodInput.txt, HGG_data.dat, tum*.dat, ROI.dat ->
   Likelihood.txt) against the oracle, and against the reference's own
   likelihood executable built into oracle/_ref when that exists.
"""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

import gliomasolver_amd as gh
from conftest import REF_DIR, ROOT
from gliomasolver_amd import FIELD_PCSF, FIELD_PG, FIELD_PHI, FIELD_PW
from gliomasolver_amd.datfile import read_dat, write_dat

pytestmark = pytest.mark.gpu
BIN = os.path.join(ROOT, "gliomasolver_amd", "bin")


def say(*a):
    print("[gpu]", *a, flush=True)


def anatomy(nx, ny, nz, seed):
    """Seeded phantom of size max(nx,ny,nz), cropped to (nz, ny, nx)."""
    m = max(nx, ny, nz)
    gm, wm, csf, _ = gh.phantom(m, seed)
    return gm[:nz, :ny, :nx].copy(), wm[:nz, :ny, :nx].copy(), csf[:nz, :ny, :nx].copy()


@pytest.mark.parametrize("real", ["f32", "f64"])
def test_ic_from_anatomy_matches_oracle(oracle, real):
    gm, wm, csf = anatomy(44, 40, 36, 3)
    ic = (0.45, 0.5, 0.4)
    say("ic_from_anatomy", real, "44x40x36 -> 32^3")
    with gh.Context(0, real) as ctx:
        ctx.grid_init_uniform(4)
        ctx.ic_from_anatomy(gm, wm, csf, ic)
        got = [ctx.get_field(f) for f in (FIELD_PW, FIELD_PG, FIELD_PCSF, FIELD_PHI)]
    exp = oracle.ic(4, gm, wm, csf, ic, real)
    for g, e in zip(got, exp):
        assert np.array_equal(g, e)
    assert (exp[3] > 0).sum() > 0 and (exp[0] > 0).sum() > 1000


def write_case(d, gm, wm, csf, ic, params):
    write_dat(os.path.join(d, "GM.dat"), gm)
    write_dat(os.path.join(d, "WM.dat"), wm)
    write_dat(os.path.join(d, "CSF.dat"), csf)
    with open(os.path.join(d, "TumorIC.txt"), "w") as f:
        f.write("%r %r %r\n" % ic)
    with open(os.path.join(d, "InputParameters.txt"), "w") as f:
        f.write("%r %r %r\n" % params)


def python_brain(gm, wm, csf, ic, params, bpd, max_level, dumpfreq, real):
    """The brain run loop (Glioma_ReactionDiffusion.cpp:272-334) through the C ABI."""
    t = np.float32 if real == "f32" else np.float64
    ic = tuple(float(t(v)) for v in ic)
    Dw, rho, tend = (t(v) for v in params)
    ctx = gh.Context(0, real)
    ctx.grid_configure(2, max_level, 0)
    ctx.grid_init_uniform(bpd)
    ctx.ic_from_anatomy(gm, wm, csf, ic)
    L = float(t(max(gm.shape) * 0.1))
    sDw, sDg, sdt, _ = gh.schedule(real, float(Dw), L, float(tend), 8 * bpd, 0.99)
    dt = t(sdt)
    ctx.refine(1e-4, 1)
    ctx.compress(1e-5, -1)
    ctx.rd_prepare(sDw, sDg, float(rho))
    tt = t(0)
    when = dumpfreq
    while tt <= tend:
        k = 0
        dump = False
        while tt <= tend:
            tt = t(tt + dt)
            k += 1
            if float(tt) >= when:
                dump = True
                break
        ctx.rd_steps(k, float(dt))
        if dump:
            ctx.refine(1e-4, 1)
            ctx.rd_prepare(sDw, sDg, float(rho))
            when = when + dumpfreq
    ctx.refine(1e-4, 1)
    out = ctx.uq_resample(8 * bpd)
    nb = ctx.nblocks()
    ctx.close()
    return out, nb


def run(cmd, cwd, timeout=300):
    say("run", " ".join(os.path.basename(c) if i == 0 else c for i, c in enumerate(cmd)))
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r.stdout


BRAIN_CASE = dict(shape=(40, 36, 44), ic=(0.45, 0.5, 0.42), params=(0.08, 0.3, 12.0), bpd=4, max_level=2,
                  dumpfreq=3.0)


@pytest.mark.parametrize("real", ["f32", "f64"])
def test_brain_uq_run_matches_python_driver(tmp_path, real):
    c = BRAIN_CASE
    gm, wm, csf = anatomy(*c["shape"][::-1], 11)
    write_case(tmp_path, gm, wm, csf, c["ic"], c["params"])
    run([os.path.join(BIN, "brain"), "-model", "RD", "-UQ", "1", "-adaptive", "1", "-vtk", "0", "-dumpfreq",
         str(c["dumpfreq"]), "-PatFileName", str(tmp_path) + "/", "-bpd", str(c["bpd"]), "-maxLevel",
         str(c["max_level"]), "-precision", "float" if real == "f32" else "double"], tmp_path)
    got = read_dat(os.path.join(tmp_path, "HGG_data.dat"))
    exp, nb = python_brain(gm, wm, csf, c["ic"], c["params"], c["bpd"], c["max_level"], c["dumpfreq"], real)
    say("brain blocks at end", nb, "max phi", float(got.max()))
    assert got.shape == (32, 32, 32)
    assert np.array_equal(got, exp)
    assert got.max() > 0 and (got > 0).sum() > 8


def test_rd_pipeline_matches_brain(tmp_path):
    c = BRAIN_CASE
    gm, wm, csf = anatomy(*c["shape"][::-1], 12)
    write_case(tmp_path, gm, wm, csf, c["ic"], c["params"])
    common = ["-model", "RD", "-UQ", "1", "-dumpfreq", str(c["dumpfreq"]), "-PatFileName", str(tmp_path) + "/",
              "-bpd", str(c["bpd"]), "-maxLevel", str(c["max_level"])]
    run([os.path.join(BIN, "brain")] + common, tmp_path)
    a = read_dat(os.path.join(tmp_path, "HGG_data.dat"))
    os.remove(os.path.join(tmp_path, "HGG_data.dat"))
    run([os.path.join(BIN, "rd_pipeline")] + common, tmp_path)
    b = read_dat(os.path.join(tmp_path, "HGG_data.dat"))
    assert np.array_equal(a, b)


def _observations(model, seed):
    rng = np.random.default_rng(seed)
    t1c = (model >= 0.7).astype(np.float64)
    flair = (model >= 0.25).astype(np.float64)
    for m in (t1c, flair):
        flip = rng.uniform(size=m.shape) < 0.05
        m[flip] = 1 - m[flip]
    pet = np.where(t1c > 0, np.clip(model + rng.normal(0, 0.05, model.shape), 0, None), 0.0)
    return t1c, flair, pet


@pytest.mark.parametrize("use_roi", [False, True])
def test_likelihood_executable(tmp_path, oracle, use_roi):
    n = 32
    rng = np.random.default_rng(5)
    z, y, x = np.meshgrid(*(np.arange(n),) * 3, indexing="ij")
    model = np.clip(1.2 - np.sqrt((x - 15.3) ** 2 + (y - 16.1) ** 2 + (z - 14.7) ** 2) / 8, 0, 1)
    t1c, flair, pet = _observations(model, 6)
    for name, a in (("HGG_data", model), ("tumT1c", t1c), ("tumFLAIR", flair), ("tumPET", pet)):
        write_dat(os.path.join(tmp_path, name + ".dat"), a)
    lik = (0.000361, 0.85, 0.7, 0.25, 2.0)
    with open(os.path.join(tmp_path, "LikelihoodInput.txt"), "w") as f:
        f.write("\n".join(repr(v) for v in lik) + "\n")
    roi = None
    args = ["-stepPET", "2", "-stepMRI", "1"]
    if use_roi:
        roi = np.unique(rng.integers(0, n, (400, 3)), axis=0).astype(np.float64)
        write_dat(os.path.join(tmp_path, "ROI.dat"), roi.T)
        args += ["-bROI", "1"]
    run([os.path.join(BIN, "likelihood")] + args, tmp_path)
    got = float(open(os.path.join(tmp_path, "Likelihood.txt")).read())
    total, _, _ = oracle.likelihood(model, pet, t1c, flair, roi=roi, stepPET=2, stepTi=1)
    say("likelihood", got, "oracle", total)
    assert abs(got - total) <= 1.5e-6            # "%Lf": 6 decimals
    ref = os.path.join(REF_DIR, "likelihood")
    if os.path.exists(ref):
        d = tmp_path / "ref"
        d.mkdir()
        for name in os.listdir(tmp_path):
            if name.endswith(".dat") or name == "LikelihoodInput.txt":
                os.symlink(os.path.join(tmp_path, name), d / name)
        run([ref] + args, d)
        assert abs(float(open(d / "Likelihood.txt").read()) - got) <= 1.5e-6
