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
