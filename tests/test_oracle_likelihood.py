"""The oracle's likelihood restatement pinned against the reference
likelihood executable (tools/pi4u_lite/Inference/Likelihood, built into
oracle/_ref from its own sources; outputs committed in
tests/golden/likelihood24.npz), plus closed forms."""
from __future__ import annotations

import math
import os

import numpy as np
import pytest

from conftest import GOLDEN

# the reference prints %Lf: 6 decimals
QUANT = 0.5e-6 + 1e-9


def golden():
    z = np.load(os.path.join(GOLDEN, "likelihood24.npz"))
    return {k: z[k] for k in z.files}


@pytest.mark.parametrize("tag", ["full", "stride", "roiA", "roiB"])
def test_oracle_likelihood_matches_reference_executable(oracle, tag):
    g = golden()
    total_ref, lpet, lt1, lt2, sp, sm = g[tag]
    roi = g["roi"] if tag.startswith("roi") else None
    total, comps, npet = oracle.likelihood(g["model"], g["pet"], g["t1c"], g["flair"], roi=roi,
                                           stepPET=int(sp), stepTi=int(sm))
    assert abs(total - total_ref) <= QUANT * 4
    for a, b in zip(comps, (lpet, lt1, lt2)):
        assert abs(a - b) <= QUANT * 2


def test_log_bernoulli_special_cases(oracle):
    """alpha = 1/2 at u = uc; constant 1/2 + 1/2(1-e^{-1/slope}) above uc;
    HGG_Likelihood.cpp:136-163."""
    n = 4
    model = np.zeros((n, n, n))
    model[0, 0, 0] = 0.7          # u == T1uc
    model[0, 0, 1] = 0.9          # u > uc
    model[0, 0, 2] = 0.95
    t1 = np.zeros_like(model)
    t1[0, 0, :3] = 1
    fl = np.zeros_like(model)
    pet = np.zeros_like(model)
    total, (lpet, lt1, lt2), npet = oracle.likelihood(model, pet, t1, fl)
    assert npet == 0 and lpet == 0.0
    a_hi = 0.5 + 0.5 * (1 - math.exp(-1 / 2.0))
    a_zero = 0.5 - 0.5 * (1 - math.exp(-(0.7 ** 2) / 2.0))
    expect_t1 = math.log(0.5) + 2 * math.log(a_hi) + (n ** 3 - 3) * math.log(1 - a_zero)
    assert lt1 == pytest.approx(expect_t1, rel=1e-14)


def test_pet_term_closed_form_and_stride(oracle):
    rng = np.random.default_rng(2)
    n = 9
    model = rng.uniform(0, 1, (n, n, n))
    pet = rng.uniform(-0.5, 1, (n, n, n))
    z = np.zeros_like(model)
    s2, b = 0.000361, 0.85
    for step in (1, 2, 3):
        _, (lpet, _, _), npet = oracle.likelihood(model, pet, z, z, stepPET=step)
        sel = np.zeros_like(model, bool)
        sel[::step, ::step, ::step] = True
        sel &= pet > 0
        sse = ((model - b * pet) ** 2)[sel].sum()
        assert npet == sel.sum()
        assert lpet == pytest.approx(-0.5 * npet * math.log(2 * math.pi * s2) - 0.5 / s2 * sse, rel=1e-12)
