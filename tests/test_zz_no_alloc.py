"""gh_likelihood_noalloc (the allocation-free variant of gh_likelihood)
through the C ABI: the same result bit for bit as gh_likelihood on two calls
in a row, the second on a smaller grid, with and without an ROI."""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np
import pytest

import gliomasolver_amd as gh
from gliomasolver_amd import FIELD_PHI
from gliomasolver_amd import workloads as wl

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu]", *a, flush=True)
    sys.stdout.flush()


def observations(u, seed):
    rng = np.random.default_rng(seed)
    t1 = (u >= 0.7).astype(np.float64)
    fl = (u >= 0.25).astype(np.float64)
    for m in (t1, fl):
        flip = rng.uniform(0, 1, u.shape) < 0.05
        m[flip] = 1 - m[flip]
    pet = np.where(t1 > 0, u + rng.normal(0, 0.05, u.shape), 0.0)
    return t1, fl, pet


def both(ctx, use_roi, params=gh.Context.LIK_DEFAULT, stepPET=1, stepTi=1):
    lib = ctx.lib
    p = (C.c_double * 5)(*params)
    res = []
    for fn in (lib.gh_likelihood, lib.gh_likelihood_noalloc):
        out = (C.c_double * 4)()
        npet = C.c_int64()
        gh.check(fn(ctx.h, p, stepPET, stepTi, int(use_roi), out, C.byref(npet)))
        res.append((np.array(out[:]), npet.value))
    return res


@pytest.mark.parametrize("real", ["f32", "f64"])
def test_noalloc_likelihood_equals_gh_likelihood(real):
    for n, tend, use_roi, steps in ((64, 150.0, False, (1, 1)), (32, 80.0, True, (2, 1))):
        case = wl.make_case(n, real, 5, Dw_phys=0.013, rho=0.1, tend=tend)
        say("noalloc likelihood", real, n, "roi" if use_roi else "full")
        ctx = wl.load_context(case)
        ctx.rd_steps(case.nsteps, case.dt)
        u = ctx.get_field(FIELD_PHI).astype(np.float64)
        t1, fl, pet = observations(u, n)
        roi = np.argwhere(fl > 0)[:, ::-1].astype(np.float64) if use_roi else None
        ctx.lik_set_data(t1, fl, pet, roi=roi)
        (a, na), (b, nb) = both(ctx, use_roi, stepPET=steps[0], stepTi=steps[1])
        assert na == nb and na > 0
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), (a, b)
        # a second evaluation on the same grid (the scratch keeps the last call's contents)
        (a2, _), (b2, _) = both(ctx, use_roi, params=(0.0025, 1.0, 0.7, 0.25, 2.0))
        assert np.array_equal(a2.view(np.uint64), b2.view(np.uint64))
        ctx.close()
