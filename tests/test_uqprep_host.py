"""`brain -model UQprep` (Glioma/Glioma_UQ_DataPreprocessing.cpp) on the host,
no GPU: observation volumes, prior ranges and the inference ROI of a
synthetic patient whose lesions are spheres, checked against what geometry
fixes -- the centre of mass of a symmetric lesion is its centre, the
equivalent radius follows from the voxel count, the ROI is the tissue inside
a sphere -- and against a brute-force restatement of the voxel mapping."""
from __future__ import annotations

import math
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from gliomasolver_amd.datfile import read_dat, write_dat

BIN = os.path.join(ROOT, "gliomasolver_amd", "bin")
# tools/scripts/GenericPrior.txt of the reference: D, rho, (T), (icx), (icy),
# (icz), PETsigma, PETscale, ucT1, ucT2, slope ranges
GENERIC = [(0.0002, 0.013), (0.0027, 0.19), (0, 0), (0, 0), (0, 0), (0, 0), (0.018, 0.25), (0.6, 1.02),
           (0.6, 0.8), (0.05, 0.6), (0.05, 0.08)]
BPD, NV = 4, 32  # grid of 32^3 cells; data of 32^3 voxels (one voxel per cell)
CENTRE, R_T1, R_FL = (15, 17, 14), 3.2, 6.4  # lesion centre (voxels) and radii


def coord(i, bpd=BPD):
    dil = np.float32(1.0 / bpd)
    h = dil / np.float32(8)
    return np.float32((i // 8) * dil) + h * np.float32(i % 8)


def patient(tmp_path):
    d = tmp_path / "pat"
    d.mkdir()
    z, y, x = np.meshgrid(np.arange(NV), np.arange(NV), np.arange(NV), indexing="ij")
    r = np.sqrt((x - CENTRE[0]) ** 2 + (y - CENTRE[1]) ** 2 + (z - CENTRE[2]) ** 2)
    brain = np.sqrt((x - 15.5) ** 2 + (y - 15.5) ** 2 + (z - 15.5) ** 2) < 13
    wm = np.where(brain, 0.7, 0.0)
    gm = np.where(brain, 0.3, 0.0)
    csf = np.where(brain, 0.0, 1.0)
    csf[(np.abs(x - 20) < 2) & (np.abs(y - 10) < 2) & brain] = 0.5  # a CSF pocket inside (CSF > 0.2 rule)
    t1 = (r <= R_T1).astype(float)
    fl = (r <= R_FL).astype(float) * 0.8
    fet = np.exp(-r ** 2 / 20.0) * 3.0
    for name, v in (("GM", gm), ("WM", wm), ("CSF", csf), ("Tum_T1c", t1), ("Tum_FLAIR", fl), ("Tum_FET", fet)):
        write_dat(str(d / (name + ".dat")), v)
    return d, dict(gm=gm, wm=wm, csf=csf, t1=t1, fl=fl, fet=fet)


def run_prep(tmp_path, pat, synt=0):
    work = tmp_path / ("work%d" % synt)
    work.mkdir()
    (work / "GenericPrior.txt").write_text("".join("%g   %g\n" % p for p in GENERIC))
    r = subprocess.run([os.path.join(BIN, "brain"), "-model", "UQprep", "-PatFileName", str(pat) + "/", "-bpd",
                        str(BPD), "-bSynt", str(synt), "-verbose", "0"], cwd=work, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr + r.stdout
    return work


def test_uqprep_observations_prior_and_roi(tmp_path):
    pat, v = patient(tmp_path)
    work = run_prep(tmp_path, pat)
    # one voxel per cell: voxel floor(coord / (1/32)) is the cell index itself
    vox = np.floor(np.array([coord(i) for i in range(NV)], np.float64) / (1.0 / NV)).astype(int)
    assert np.array_equal(vox, np.arange(NV))
    t1 = read_dat(str(work / "tumT1c.dat"))
    fl = read_dat(str(work / "tumFLAIR.dat"))
    pet = read_dat(str(work / "tumPET.dat"))
    assert np.array_equal(t1, (v["t1"] > 0.1).astype(float))
    assert np.array_equal(fl, (v["fl"] > 0.1).astype(float))
    assert pet.max() == 1.0  # normalised by the maximum
    assert np.allclose(pet, v["fet"].astype(np.float32) / np.float32(v["fet"].max()), rtol=1e-6, atol=0)

    # prior ranges: generic rows kept, IC box = centre -/+ 1.25 r_T1, T range from the FLAIR volume
    pr = np.loadtxt(str(work / "PriorRange.txt"))
    lpr = np.loadtxt(str(work / "LogPriorRange.txt"))
    assert pr.shape == (11, 2)
    for k in (0, 1, 6, 7, 8, 9, 10):
        assert pr[k] == pytest.approx(GENERIC[k], abs=1e-4)
    h = 1.0 / (8 * BPD)
    n1, n2 = int(t1.sum()), int(fl.sum())
    r1 = 1.25 * (3 * n1 * h ** 3 / (4 * math.pi)) ** (1 / 3)
    centre = [coord(c) for c in CENTRE]  # the lesion is symmetric about a cell: centre of mass = that cell
    for a in range(3):
        assert pr[3 + a] == pytest.approx([centre[a] - r1, centre[a] + r1], abs=1.5e-4)
    L = NV * 0.1
    rfl = 1.25 * (3 * n2 * h ** 3 / (4 * math.pi)) ** (1 / 3) * L
    vmin, vmax = 2 * math.sqrt(GENERIC[0][0] * GENERIC[1][0]), 2 * math.sqrt(GENERIC[0][1] * GENERIC[1][1])
    tmin, tmax = max(30.0, rfl / vmax), min(1800.0, (rfl + 2) / vmin + 300)
    assert pr[2] == pytest.approx([tmin * tmin * GENERIC[1][1], tmax * tmax * GENERIC[1][0]], rel=1e-5, abs=1e-4)
    with np.errstate(divide="ignore"):
        assert np.allclose(lpr[pr > 0], np.log(pr[pr > 0]), atol=2e-4)
    assert np.all(np.isneginf(lpr[pr == 0]) | (lpr[pr == 0] < -1e30))

    # ROI: cells with tissue (p_w + p_g > 0) within twice the FLAIR-equivalent radius
    roi = read_dat(str(work / "ROI.dat"))  # (3, N)
    rr = 2 * (3 * n2 * h ** 3 / (4 * math.pi)) ** (1 / 3)
    c = np.array([coord(i) for i in range(NV)], np.float64)
    X, Y, Z = np.meshgrid(c, c, c, indexing="xy")
    Z = np.broadcast_to(c[:, None, None], (NV, NV, NV))
    Y = np.broadcast_to(c[None, :, None], (NV, NV, NV))
    X = np.broadcast_to(c[None, None, :], (NV, NV, NV))
    dist = np.sqrt((X - centre[0]) ** 2 + (Y - centre[1]) ** 2 + (Z - centre[2]) ** 2)
    tissue = (v["wm"] + v["gm"] > 0) & ~((v["csf"] > 0.2) & (v["csf"] < 1))
    inside = tissue & (dist <= rr)
    margin = np.abs(dist - rr) < 1e-5  # cells on the sphere itself may fall either way in float
    got = np.zeros((NV, NV, NV), bool)
    got[roi[2].astype(int), roi[1].astype(int), roi[0].astype(int)] = True
    assert roi.shape[1] == int(got.sum())  # no duplicates
    assert np.array_equal(got & ~margin, inside & ~margin)
    # listed in block order: block (z, y, x), then cells (z, y, x)
    key = [(int(k) // 8, int(j) // 8, int(i) // 8, int(k) % 8, int(j) % 8, int(i) % 8) for i, j, k in roi.T]
    assert key == sorted(key)


def test_uqprep_synthetic_observations(tmp_path):
    """-bSynt 1: masks of a ground-truth density (> 0.7, > 0.25), PET = the
    density where either mask holds plus N(0, 0.02^2) noise, deterministic
    (drand48's default seed)."""
    d = tmp_path / "gt"
    d.mkdir()
    z, y, x = np.meshgrid(np.arange(NV), np.arange(NV), np.arange(NV), indexing="ij")
    r = np.sqrt((x - CENTRE[0]) ** 2 + (y - CENTRE[1]) ** 2 + (z - CENTRE[2]) ** 2)
    gt = np.clip(1.0 - r / 9.0, 0, 1)
    write_dat(str(d / "HGG_data_GT.dat"), gt)
    w1 = run_prep(tmp_path, d, synt=1)
    t1 = read_dat(str(w1 / "tumT1c.dat"))
    fl = read_dat(str(w1 / "tumFLAIR.dat"))
    pet = read_dat(str(w1 / "tumPET.dat"))
    g32 = gt.astype(np.float32)
    assert np.array_equal(t1, (g32 > np.float32(0.7)).astype(float))
    assert np.array_equal(fl, (g32 > np.float32(0.25)).astype(float))
    assert np.all(pet[fl == 0] == 0) and pet.max() == 1.0  # PET where either mask holds
    m = fl == 1
    k = np.median(pet[m] / g32[m])  # the normalisation 1 / max(gt + noise)
    resid = pet[m] - k * g32[m]
    assert 0.012 * k < np.std(resid) < 0.03 * k and abs(np.mean(resid)) < 0.005 * k
    # no tissue in the synthetic case, so the ROI is empty (as in the reference)
    assert read_dat(str(w1 / "ROI.dat")).shape[1] == 0
    (tmp_path / "again").mkdir()
    w2 = run_prep(tmp_path / "again", d, synt=1)
    assert np.array_equal(read_dat(str(w2 / "tumPET.dat")), pet)
