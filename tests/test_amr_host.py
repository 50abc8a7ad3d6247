"""Wavelet-adaptive grid on the CPU: the library's host plan logic against a
brute-force restatement of MRAG's rules, and the oracle's ghost / FWT / split /
multi-level RD restatement pinned by closed forms."""
from __future__ import annotations

import numpy as np
import pytest

import gliomasolver_amd as gh
from amr_util import (apply_refine, balanced_random_grid, canonical, cell_positions, collapse_parents, neighbours,
                      neighbours_by_ring, refine_victims)

SEEDS = [0, 1, 2, 3]


@pytest.mark.parametrize("seed", SEEDS)
def test_neighbourhood_matches_brute_force(seed):
    rows = balanced_random_grid(np.random.default_rng(seed), base_level=1, max_level=3, nrefine=3)
    ref = neighbours(rows)
    for leaf in range(len(rows)):
        assert gh.plan_neighbours(rows, leaf) == ref[leaf], leaf


@pytest.mark.parametrize("base,top", [(0, 2), (1, 3), (2, 4)])
def test_ring_neighbourhood_equals_brute_force(base, top):
    """The linear-time neighbourhood used by the plan restatements on large
    grids gives the brute-force box test's (leaf, periodic image) pairs,
    including the tiny periodic grids where a leaf touches itself."""
    for seed in range(3):
        rows = balanced_random_grid(np.random.default_rng(40 + seed), base_level=base, max_level=top, nrefine=3)
        assert neighbours_by_ring(rows) == neighbours(rows)


@pytest.mark.parametrize("seed", SEEDS)
@pytest.mark.parametrize("max_level", [2, 3, 4])
def test_refinement_plan_matches_restatement(seed, max_level):
    rng = np.random.default_rng(100 + seed)
    rows = balanced_random_grid(rng, base_level=1, max_level=3, nrefine=3)
    flags = (rng.uniform(0, 1, len(rows)) < 0.2).astype(np.uint8)
    got = gh.plan_refine(rows, flags, max_level)
    assert got == refine_victims(rows, flags, max_level)
    new = apply_refine(rows, got)
    nb = neighbours(new)
    assert all(abs(int(new[s, 0]) - int(new[j, 0])) <= 1 for s in range(len(new)) for j, _ in nb[s])
    assert new[:, 0].max() <= max(max_level, rows[:, 0].max())


@pytest.mark.parametrize("seed", SEEDS)
@pytest.mark.parametrize("min_level", [0, 2])
def test_compression_plan_matches_restatement(seed, min_level):
    rng = np.random.default_rng(200 + seed)
    rows = balanced_random_grid(rng, base_level=1, max_level=3, nrefine=4)
    flags = (rng.uniform(0, 1, len(rows)) < 0.8).astype(np.uint8)
    got = gh.plan_compress(rows, flags, min_level)
    assert got == collapse_parents(rows, flags, min_level)


@pytest.mark.parametrize("seed", range(4))
def test_compression_neighbourhood_rule_on_every_parent(seed):
    """All leaves flagged, so the plan is every parent of 8 leaves whose
    children's neighbours lie within one level of it: the neighbourhood rule
    alone, on grids of levels 2-5 (parents at levels 1-4, periodic wrap)."""
    rng = np.random.default_rng(300 + seed)
    rows = canonical([[2, x, y, z] for z in range(4) for y in range(4) for x in range(4)])
    pts = rng.uniform(0, 1, (3, 3))
    for L in range(2, 5):  # refine around three points down to level 5, keeping 2:1 balance
        rows = apply_refine(rows, [k for k in range(len(rows)) if rows[k, 0] == L and
                                   any(all(int(p[a] * (1 << L)) == rows[k, 1 + a] for a in range(3)) for p in pts)])
        while True:
            nb = neighbours_by_ring(rows)
            bad = sorted({j for s in range(len(rows)) for j, _ in nb[s] if rows[s, 0] - rows[j, 0] > 1})
            if not bad:
                break
            rows = apply_refine(rows, bad)
    assert rows[:, 0].max() == 5
    flags = np.ones(len(rows), np.uint8)
    got = gh.plan_compress(rows, flags, 0)
    assert got == collapse_parents(rows, flags, 0)
    assert got  # some parents pass, and (below) some fail the rule
    allp = {(L - 1, x >> 1, y >> 1, z >> 1) for L, x, y, z in rows.tolist() if L > 0}
    key = {tuple(r) for r in rows.tolist()}
    full = [p for p in allp if all((p[0] + 1, 2 * p[1] + (o & 1), 2 * p[2] + ((o >> 1) & 1), 2 * p[3] + (o >> 2))
                                   in key for o in range(8))]
    assert len(full) > len(got)


def test_plans_on_a_grid_deeper_than_the_scanned_maps():
    """Levels 2-8 around one point (2:1-balanced): the plans read the small
    levels' maps directly and the deep levels (dense 6-7, hashed 8) through
    the general lookup; both plans equal the restatement."""
    rows = canonical([[1, x, y, z] for z in range(2) for y in range(2) for x in range(2)])
    p = (0.3, 0.45, 0.6)
    for L in range(1, 8):
        rows = apply_refine(rows, [k for k in range(len(rows)) if rows[k, 0] == L and
                                   all(int(p[a] * (1 << L)) == rows[k, 1 + a] for a in range(3))])
        while True:
            nb = neighbours_by_ring(rows)
            bad = sorted({j for s in range(len(rows)) for j, _ in nb[s] if rows[s, 0] - rows[j, 0] > 1})
            if not bad:
                break
            rows = apply_refine(rows, bad)
    assert rows[:, 0].max() == 8
    rng = np.random.default_rng(3)
    flags = (rng.uniform(0, 1, len(rows)) < 0.2).astype(np.uint8)
    assert gh.plan_refine(rows, flags, 9) == refine_victims(rows, flags, 9)
    flags = (rng.uniform(0, 1, len(rows)) < 0.8).astype(np.uint8)
    assert gh.plan_compress(rows, flags, 0) == collapse_parents(rows, flags, 0)


def test_plans_reject_nested_and_repeated_leaves():
    """A leaf inside another leaf, or a leaf given twice, is not a tree: the
    plans refuse it, on the dense levels and on the hashed ones (levels > 7)."""
    from gliomasolver_amd._capi import GHError
    rows = canonical([[1, x, y, z] for z in range(2) for y in range(2) for x in range(2)])
    zero = lambda r: np.zeros(len(r), np.uint8)
    cases = [
        canonical(np.vstack([rows, [[2, 0, 0, 0]]])),          # level-2 leaf inside a level-1 leaf
        canonical(np.vstack([rows, [[8, 1, 1, 1]]])),          # hashed leaf inside a dense leaf
        canonical(np.vstack([rows, [[1, 0, 0, 0]]])),          # repeated dense leaf
    ]
    deep = canonical(np.vstack([rows[1:], [[8, 0, 0, 0], [8, 0, 0, 0]]]))  # repeated hashed leaf
    cases.append(deep)
    deep2 = canonical(np.vstack([rows[1:], [[8, 0, 0, 0], [9, 0, 0, 1]]]))  # hashed leaf inside a hashed leaf
    cases.append(deep2)
    for r in cases:
        with pytest.raises(GHError, match="inside another leaf"):
            gh.plan_refine(r, zero(r), 9)
    assert gh.plan_refine(rows, zero(rows), 3) == []


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("side", ["low", "high"])
def test_chain_reaction_across_the_periodic_wrap(axis, side):
    """Half of a level-2 grid refined to level 3 along one axis; one level-3
    block on the periodic boundary is flagged.  Its only coarser neighbour
    lies across the wrap, where the space extension does not look, so only
    the 2:1 chain reaction (MRAGRefiner.h:65-152) refines it."""
    rows = canonical([[2, x, y, z] for z in range(4) for y in range(4) for x in range(4)])
    half = {0, 1} if side == "low" else {2, 3}
    rows = apply_refine(rows, [k for k in range(len(rows)) if rows[k, 1 + axis] in half])
    idx = [3, 3, 3] if side == "low" else [4, 4, 4]
    idx[axis] = 0 if side == "low" else 7
    v = next(k for k in range(len(rows)) if rows[k, 0] == 3 and list(rows[k, 1:]) == idx)
    flags = np.zeros(len(rows), np.uint8)
    flags[v] = 1
    got = gh.plan_refine(rows, flags, 5)
    assert got == refine_victims(rows, flags, 5)
    across = [i // 2 for i in idx]
    across[axis] = 3 if side == "low" else 0
    w = next(k for k in range(len(rows)) if rows[k, 0] == 2 and list(rows[k, 1:]) == across)
    assert v in got and w in got


def test_space_extension_refines_coarse_neighbours_of_fine_blocks():
    """With no flags at all, a leaf next to a finer leaf is still refined
    (MRAGRefiner_SpaceExtension.h:35-51) -- except across the periodic wrap."""
    rows = canonical([[1, x, y, z] for z in range(2) for y in range(2) for x in range(2)])
    rows = apply_refine(rows, [0])  # refine block (0,0,0) at level 1
    victims = gh.plan_refine(rows, np.zeros(len(rows), np.uint8), 3)
    lv = rows[victims]
    assert len(victims) == 7 and np.all(lv[:, 0] == 1)


# ---------------------------------------------------------------------------
# oracle pins
# ---------------------------------------------------------------------------

def field_on(rows, f, dtype=np.float64):
    pos = cell_positions(rows)
    return f(pos[..., 0], pos[..., 1], pos[..., 2]).astype(dtype).reshape(len(rows), 512)


def central_two_level_grid():
    """level-3 grid with its central 2x2x2 blocks refined to level 4 (balanced)"""
    rows = canonical([[3, x, y, z] for z in range(8) for y in range(8) for x in range(8)])
    centre = [i for i, (L, x, y, z) in enumerate(rows.tolist()) if 3 <= x <= 4 and 3 <= y <= 4 and 3 <= z <= 4]
    return apply_refine(rows, centre)


def interior_leaves(rows, margin_cells=10):
    """leaves whose lab plus interpolation support stays inside the unit cube"""
    out = []
    for b, (L, x, y, z) in enumerate(canonical(rows).tolist()):
        h = 2.0 ** -L / 8
        lo = np.array([x, y, z]) * 2.0 ** -L
        if np.all(lo - margin_cells * h > 0) and np.all(lo + (8 + margin_cells) * h < 1):
            out.append(b)
    return out


def lab_positions(rows, b, order):
    L, x, y, z = canonical(rows)[b].tolist()
    lo, n = (-2, 13) if order == 4 else (-1, 10)
    i = np.arange(n) + lo
    h = 2.0 ** -L / 8
    zz, yy, xx = np.meshgrid(i, i, i, indexing="ij")
    return x * 2.0 ** -L + h * xx, y * 2.0 ** -L + h * yy, z * 2.0 ** -L + h * zz


@pytest.mark.parametrize("order,deg_ok,deg_bad", [(4, 3, 4), (2, 1, 2)])
def test_oracle_ghosts_reproduce_polynomials(oracle, order, deg_ok, deg_bad):
    rows = central_two_level_grid()
    assert rows[:, 0].min() < rows[:, 0].max()
    ok = lambda x, y, z: 0.3 + 0.7 * (x - 0.41) ** deg_ok - 0.5 * (y - 0.52) ** deg_ok * (z - 0.47)
    bad = lambda x, y, z: (x - 0.41) ** deg_bad + (y - 0.6) ** deg_bad + (z - 0.45) ** deg_bad
    f_ok, f_bad = field_on(rows, ok), field_on(rows, bad)
    leaves = interior_leaves(rows)
    assert len(leaves) > 4
    worst_bad = 0.0
    for b in leaves:
        X, Y, Z = lab_positions(rows, b, order)
        np.testing.assert_allclose(oracle.amr_lab(rows, order, f_ok, b), ok(X, Y, Z), rtol=0, atol=1e-13)
        worst_bad = max(worst_bad, np.abs(oracle.amr_lab(rows, order, f_bad, b) - bad(X, Y, Z)).max())
    assert worst_bad > 1e-9  # the next degree is not reproduced across a level jump


@pytest.mark.parametrize("order,deg", [(4, 3), (2, 1)])
def test_oracle_fwt_vanishing_moments_and_impulse(oracle, order, deg):
    rows = canonical([[2, x, y, z] for z in range(4) for y in range(4) for x in range(4)])
    f = field_on(rows, lambda x, y, z: 0.2 + (x - 0.3) ** deg * (1 + y) + (z - 0.7) ** deg)
    g = field_on(rows, lambda x, y, z: np.sin(7 * x) * np.cos(5 * y) * np.sin(3 * z + 1))
    b = 21  # block (1,1,1): interior
    assert oracle.amr_fwt(order, oracle.amr_lab(rows, order, f, b)) < 1e-12
    assert oracle.amr_fwt(order, oracle.amr_lab(rows, order, g, b)) > 1e-4
    imp = np.zeros((len(rows), 512))
    imp[b, 3 + 8 * 2 + 64 * 4] = 1.0  # x odd, y and z even: the x-wavelet's centre tap Ga = 1
    assert oracle.amr_fwt(order, oracle.amr_lab(rows, order, imp, b)) == 1.0


@pytest.mark.parametrize("order,rate", [(4, 16.0), (2, 4.0)])
def test_oracle_split_convergence_order(oracle, order, rate):
    """One refinement of a periodic smooth field: the L-inf error shrinks by
    2^order per halving of h (the convergence pin of
    MRAG/MRAGcore/MRAG_STDTestL2_Refinement.h:121-133)."""
    f = lambda x, y, z: np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y) + 0.3 * np.sin(2 * np.pi * z)
    errs = []
    for L in (1, 2, 3):
        n = 1 << L
        rows = canonical([[L, x, y, z] for z in range(n) for y in range(n) for x in range(n)])
        vals = field_on(rows, f)
        e = 0.0
        for b in range(len(rows)):
            lab = oracle.amr_lab(rows, order, vals, b)
            Lb, x, y, z = rows[b].tolist()
            for o in range(8):
                ch = oracle.amr_split(order, lab, o)
                cx, cy, cz = 2 * x + (o & 1), 2 * y + ((o >> 1) & 1), 2 * z + ((o >> 2) & 1)
                pos = cell_positions([[L + 1, cx, cy, cz]])[0]
                e = max(e, np.abs(ch - f(pos[..., 0], pos[..., 1], pos[..., 2])).max())
        errs.append(e)
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert abs(np.log2(r2) - np.log2(rate)) < 0.25, errs


def test_oracle_split_reproduces_cubics_exactly(oracle):
    rows = canonical([[2, x, y, z] for z in range(4) for y in range(4) for x in range(4)])
    f = lambda x, y, z: 0.1 + (x - 0.5) ** 3 + (y - 0.4) ** 2 * (z - 0.6)
    vals = field_on(rows, f)
    b = 21
    lab = oracle.amr_lab(rows, 4, vals, b)
    for o in range(8):
        ch = oracle.amr_split(4, lab, o)
        pos = cell_positions([[3, 2 + (o & 1), 2 + ((o >> 1) & 1), 2 + ((o >> 2) & 1)]])[0]
        np.testing.assert_allclose(ch, f(pos[..., 0], pos[..., 1], pos[..., 2]), rtol=0, atol=1e-14)


@pytest.mark.parametrize("real", ["f32", "f64"])
def test_oracle_multilevel_rd_reduces_to_uniform_oracle(oracle, real):
    """On a one-level grid the multi-level restatement equals the uniform one
    (itself bit-exact to the reference's golden vectors)."""
    t = np.float32 if real == "f32" else np.float64
    rng = np.random.default_rng(5)
    n = 16
    pw = rng.uniform(0, 1, (n, n, n)).astype(t)
    pg = ((1 - pw) * rng.uniform(0, 1, pw.shape)).astype(t)
    pw[rng.uniform(0, 1, pw.shape) < 0.15] = 0
    phi = (rng.uniform(0, 1, pw.shape) ** 3).astype(t)
    rows = canonical([[1, x, y, z] for z in range(2) for y in range(2) for x in range(2)])

    def to_blocks(a):
        return np.ascontiguousarray(a.reshape(2, 8, 2, 8, 2, 8).transpose(0, 2, 4, 1, 3, 5).reshape(8, 512))

    out = oracle.amr_rd_steps(rows, 4, to_blocks(phi), to_blocks(pw), to_blocks(pg), 2e-4, 2e-5, 0.05, 0.3, 6)
    ref = oracle.rd_steps(phi, pw, pg, 2e-4, 2e-5, 0.05, 1.0 / n, 0.3, 6)
    assert np.array_equal(out.view(np.uint8), to_blocks(ref).view(np.uint8))


def test_oracle_multilevel_rd_keeps_constant_state(oracle):
    rows = balanced_random_grid(np.random.default_rng(9), base_level=1, max_level=3, nrefine=2)
    nb = len(rows)
    phi = np.full((nb, 512), 0.5)
    pw = np.full((nb, 512), 0.75)
    pg = np.full((nb, 512), 0.25)
    out = oracle.amr_rd_steps(rows, 4, phi, pw, pg, 1e-4, 1e-5, 0.0, 0.1, 3)
    assert np.array_equal(out, phi)


def test_oracle_resample_uniform_is_identity_and_coarse_replicates(oracle):
    rows = canonical([[1, x, y, z] for z in range(2) for y in range(2) for x in range(2)])
    vals = np.arange(8 * 512, dtype=np.float64).reshape(8, 512)
    out = oracle.amr_resample(rows, vals, 16)
    dense = vals.reshape(2, 2, 2, 8, 8, 8).transpose(0, 3, 1, 4, 2, 5).reshape(16, 16, 16)
    assert np.array_equal(out, dense)
    out2 = oracle.amr_resample(rows, vals, 32)  # every cell covers 2^3 finest cells
    assert np.array_equal(out2[::2, ::2, ::2], dense) and np.array_equal(out2[1::2, 1::2, 1::2], dense)


def _far_from(rows_b, parents, top):
    """Block (L, x, y, z) dilated by two of its own block widths (periodic)
    shares no point with any parent box."""
    L = int(rows_b[0])
    n = 1 << top
    s = 1 << (top - L)
    for P in parents:
        sp = 1 << (top - int(P[0]))
        hit = True
        for a in range(3):
            if 5 * s >= n:          # the dilated interval covers the whole axis
                continue
            lo = ((int(rows_b[1 + a]) - 2) * s) % n
            hi = lo + 5 * s         # [lo, hi) at level top, possibly past n (wrap)
            plo, phi = int(P[1 + a]) * sp, (int(P[1 + a]) + 1) * sp
            if not any(max(lo - k * n, plo) < min(hi - k * n, phi) for k in (0, 1)):
                hit = False
                break
        if hit:
            return False
    return True


@pytest.mark.parametrize("order", [4, 2])
def test_lab_unchanged_two_blocks_away_from_a_collapse(oracle, order):
    """The lab of a block -- and so its FWT details -- depends only on the
    leaves within two of its block widths: a collapse whose parent boxes stay
    outside that reach leaves the lab bit-identical (the argument behind the
    compression pass reusing details between its loops: lab [-2, 11) and at
    most two synthesis levels of 3 and 6 cells reach 12 < 16 cells)."""
    from amr_util import apply_collapse
    rng = np.random.default_rng(17 + order)
    far_checked = near_changed = 0
    for trial in range(6):
        rows = balanced_random_grid(rng, base_level=3, max_level=5, nrefine=4)
        top = int(rows[:, 0].max())
        cand = collapse_parents(rows, np.ones(len(rows), np.uint8), 0)
        if not cand:
            continue
        pick = [cand[i] for i in sorted(rng.choice(len(cand), size=min(2, len(cand)), replace=False))]
        new = apply_collapse(rows, pick)
        old_index = {tuple(r): i for i, r in enumerate(rows.tolist())}
        phi_old = rng.uniform(0, 1, (len(rows), 512)).astype(np.float32)
        phi_new = np.zeros((len(new), 512), np.float32)
        for j, r in enumerate(new.tolist()):
            if tuple(r) in old_index:
                phi_new[j] = phi_old[old_index[tuple(r)]]
            else:                                   # a collapsed parent: the coincident child vertices
                L, px, py, pz = r
                c = np.arange(512)
                cx, cy, cz = c & 7, (c >> 3) & 7, c >> 6
                octant = (cx >> 2) | ((cy >> 2) << 1) | ((cz >> 2) << 2)
                src = 2 * (cx & 3) + 8 * 2 * (cy & 3) + 64 * 2 * (cz & 3)
                for o in range(8):
                    k = old_index[(L + 1, 2 * px + (o & 1), 2 * py + ((o >> 1) & 1), 2 * pz + ((o >> 2) & 1))]
                    m = octant == o
                    phi_new[j, m] = phi_old[k, src[m]]
        for j, r in enumerate(new.tolist()):
            if tuple(r) not in old_index:
                continue
            a = oracle.amr_lab(rows, order, phi_old, old_index[tuple(r)])
            b = oracle.amr_lab(new, order, phi_new, j)
            if _far_from(r, pick, top):
                far_checked += 1
                assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (trial, r, pick)
            elif not np.array_equal(a, b):
                near_changed += 1
    assert far_checked > 200 and near_changed > 10, (far_checked, near_changed)
