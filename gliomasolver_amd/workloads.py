"""Synthetic workloads: seeded brain phantoms, initial conditions and
parameter draws of the benchmark configurations.

Uniform 64^3, FP64: Dw = 0.013 cm^2/day, rho = 0.025/day, T = 300 days, CFL 0.9.
Uniform 128^3, FP32: Dw ~ logU[0.002, 0.015], rho ~ logU[0.002, 0.2], T = 400 days.
Both: a 17.6 cm cube (L = 17.6, Glioma_ReactionDiffusion.cpp:97,302), Gaussian
IC (sigma = 1 cell, peak 0.1) at a seeded WM voxel.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import Context, FIELD_PHI, gaussian_ic, phantom, schedule

L_CUBE = 17.6  # cm


@dataclass
class Case:
    n: int
    real: str
    Dw_phys: float
    rho: float
    tend: float
    cfl: float
    seed: int
    gm: np.ndarray
    wm: np.ndarray
    csf: np.ndarray
    seed_voxel: tuple
    Dw: float = 0.0
    Dg: float = 0.0
    dt: float = 0.0
    nsteps: int = 0

    @property
    def ncells(self) -> int:
        return self.n ** 3

    def describe(self) -> dict:
        return {"n": self.n, "real": self.real, "Dw_phys": self.Dw_phys, "rho": self.rho, "T_days": self.tend,
                "cfl": self.cfl, "seed": self.seed, "euler_steps": self.nsteps, "dt_days": self.dt,
                "seed_voxel": list(self.seed_voxel)}


def log_uniform(rng, lo, hi):
    return math.exp(rng.uniform(math.log(lo), math.log(hi)))


def make_case(n: int, real: str, seed: int, Dw_phys=None, rho=None, tend=None, cfl=0.9) -> Case:
    rng = np.random.default_rng(1000 * n + seed)
    if Dw_phys is None:
        Dw_phys = log_uniform(rng, 0.002, 0.015)
    if rho is None:
        rho = log_uniform(rng, 0.002, 0.2)
    if tend is None:
        tend = 400.0
    gm, wm, csf, sv = phantom(n, seed)
    c = Case(n, real, Dw_phys, rho, tend, cfl, seed, gm, wm, csf, sv)
    c.Dw, c.Dg, c.dt, c.nsteps = schedule(real, Dw_phys, L_CUBE, tend, n, cfl)
    return c


def config_uniform64(seed: int = 1) -> Case:
    return make_case(64, "f64", seed, Dw_phys=0.013, rho=0.025, tend=300.0, cfl=0.9)


def config_uniform128(seed: int = 2) -> Case:
    return make_case(128, "f32", seed, tend=400.0, cfl=0.9)


def load_context(case: Case, device: int = 0) -> Context:
    """Context with the case's anatomy (reference _ic processing) and IC uploaded, RD prepared."""
    ctx = Context(device, case.real)
    ctx.grid_init_uniform(case.n // 8)
    ctx.set_anatomy(case.gm, case.wm, case.csf)
    from . import FIELD_PG, FIELD_PW
    pw = ctx.get_field(FIELD_PW)
    pg = ctx.get_field(FIELD_PG)
    ctx.set_field(FIELD_PHI, gaussian_ic(case.real, pw, pg, case.seed_voxel, 1.0, 0.1))
    ctx.rd_prepare(case.Dw, case.Dg, case.rho)
    return ctx


@dataclass
class Stage:
    """One TMCMC stage of samples over several grid sizes (see tmcmc_stage)."""
    sizes: list           # grid size per anatomy
    anatomies: list       # (gm, wm, csf) per anatomy, data at the grid size
    wm_voxels: list       # candidate seed voxels per anatomy (WM > 0.9)
    anatomy: np.ndarray   # per sample
    params: np.ndarray    # per sample {Dw / L^2, rho, T, x0, y0, z0}
    nsteps: np.ndarray    # per sample, dt = T / N below the CFL limit 0.9 h^2 / (6 Dw)
    Dw_phys: np.ndarray
    seed: int

    @property
    def cells(self) -> np.ndarray:
        return np.array([self.sizes[a] ** 3 for a in self.anatomy], np.int64)

    def cell_steps(self) -> int:
        return int((self.cells * self.nsteps).sum())


def tmcmc_stage(nsamples: int = 1024, sizes=(64, 96, 128), weights=(1, 2, 1), seed: int = 5) -> Stage:
    """Samples of (Dw, rho, T, x0, y0, z0): Dw ~ logU[0.002, 0.015] cm^2/day,
    rho ~ logU[0.002, 0.2]/day, T ~ U[100, 700] days, the Gaussian seed at a
    uniformly drawn WM voxel of a phantom of the sample's size; sizes drawn
    in the given ratio.  dt = T/N with N the smallest count that keeps
    dt <= 0.9 h^2 / (6 Dw / L^2)."""
    rng = np.random.default_rng(5000 + seed)
    anats, wmv = [], []
    for n in sizes:
        gm, wm, csf, _ = phantom(n, seed)
        anats.append((gm, wm, csf))
        wmv.append(np.argwhere(wm > 0.9)[:, ::-1])  # (x, y, z)
    w = np.asarray(weights, np.float64)
    anatomy = rng.choice(len(sizes), size=nsamples, p=w / w.sum()).astype(np.int32)
    Dw_phys = np.exp(rng.uniform(math.log(0.002), math.log(0.015), nsamples))
    rho = np.exp(rng.uniform(math.log(0.002), math.log(0.2), nsamples))
    T = rng.uniform(100.0, 700.0, nsamples)
    params = np.zeros((nsamples, 6))
    nsteps = np.zeros(nsamples, np.int64)
    for i in range(nsamples):
        a = anatomy[i]
        n = sizes[a]
        Dw = Dw_phys[i] / (L_CUBE * L_CUBE)
        dt_max = 0.9 * (1.0 / n) ** 2 / (6.0 * Dw)
        nsteps[i] = int(math.ceil(T[i] / dt_max))
        v = wmv[a][rng.integers(len(wmv[a]))]
        params[i] = (Dw, rho[i], T[i], float(v[0]), float(v[1]), float(v[2]))
    return Stage(list(sizes), anats, wmv, anatomy, params, nsteps, Dw_phys, seed)
