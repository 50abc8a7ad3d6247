"""The ghost accumulation order of the adaptive lab is MRAG's.

MRAG evaluates a lab ghost from an instruction list built per block by
MRAG_BBInfoCreator -- easy ghosts scattered from one neighbour in its (z, y, x)
point order, the rest resolved level by level as "bastard" ghosts and
collected depth first -- and sums value * ((Real)wx*(Real)wy*(Real)wz) in Real
in list order (MRAG/MRAGcore/MRAGBlockLab.h:256-295).  The build (device
tables, device recursion and the oracle alike) sums the recursion over the
multiresolution instead.  oracle/gh_oracle_mrag.c restates MRAG's
construction on its own terms; these tests find its lists -- sources, order
and Real weights -- and the FP32/FP64 labs they give equal to the recursion's,
on random 2:1-balanced grids with both wavelet orders and every leaf.  A ghost
summed in another order would differ in its last bits on the random fields
used here."""
from __future__ import annotations

import numpy as np
import pytest

from amr_util import balanced_random_grid


def grids():
    out = []
    for seed, (base, top, nref) in enumerate([(1, 3, 3), (1, 3, 5), (2, 4, 3), (2, 4, 6), (1, 4, 4), (0, 3, 3)]):
        out.append((seed, balanced_random_grid(np.random.default_rng(100 + seed), base, top, nref)))
    return out


GRIDS = grids()


@pytest.mark.parametrize("order", [2, 4])
@pytest.mark.parametrize("real", ["f32", "f64"])
@pytest.mark.parametrize("gi", range(len(GRIDS)))
def test_mrag_instruction_lists_give_the_recursion_labs(oracle, order, real, gi):
    seed, rows = GRIDS[gi]
    assert len(set(rows[:, 0].tolist())) > 1 or rows[0, 0] == 0
    rng = np.random.default_rng(seed)
    t = np.float32 if real == "f32" else np.float64
    # values of mixed sign and spread exponents: a reordered sum changes bits
    field = (rng.uniform(-1, 1, len(rows) * 512) * np.exp(rng.uniform(-6, 6, len(rows) * 512))).astype(t)
    for leaf in range(len(rows)):
        a = oracle.amr_lab(rows, order, field, leaf)
        b = oracle.mrag_lab(rows, order, field, leaf)
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), (seed, leaf, rows[leaf].tolist())


@pytest.mark.parametrize("order", [2, 4])
@pytest.mark.parametrize("f32", [True, False])
def test_mrag_instruction_lists_equal_recursion_terms(oracle, order, f32):
    """Term by term: the same sources in the same order with the same Real
    weight, for every lab ghost of every leaf."""
    ghosts = multi = 0
    for seed, rows in GRIDS:
        for leaf in range(len(rows)):
            ca, sa, wa = oracle.block_terms("amr", rows, order, leaf, f32)
            cb, sb, wb = oracle.block_terms("mrag", rows, order, leaf, f32)
            assert np.array_equal(ca, cb), (seed, leaf)
            assert np.array_equal(sa, sb) and np.array_equal(wa, wb), (seed, leaf)
            ghosts += int((ca > 0).sum())
            multi += int((ca > 1).sum())
    assert ghosts > 50000 and multi > 5000, (ghosts, multi)


def test_mrag_restatement_refuses_unbalanced_grid(oracle):
    """A neighbour two levels away is outside what MRAG's 2:1 grids hold."""
    rows = np.array([[1, 0, 0, 0], [1, 1, 0, 0], [1, 0, 1, 0], [1, 1, 1, 0], [1, 0, 0, 1], [1, 1, 0, 1], [1, 0, 1, 1]]
                    + [[3, 4 + (o & 1) * 2 + x, 4 + ((o >> 1) & 1) * 2 + y, 4 + ((o >> 2) & 1) * 2 + z]
                       for o in range(8) for z in range(2) for y in range(2) for x in range(2)], np.int32)
    from amr_util import canonical
    rows = canonical(rows)
    f = np.zeros(len(rows) * 512, np.float32)
    r = rows
    lab = np.zeros(13 ** 3, np.float32)
    rc = oracle.lib.gho_mrag_lab_f32(len(r), oracle.p(np.ascontiguousarray(r)), 4, oracle.p(f), 0, oracle.p(lab))
    assert rc == -1
