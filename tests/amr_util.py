"""Test-side restatement of MRAG's block-neighbourhood and plan rules by
brute-force box geometry (checker code, independent of the library's
direction walk):

  neighbourhood  MRAG/MRAGcore/MRAGrid.inl:840-1149, MRAGridHelpers.h:77-84 --
                 leaves whose finest-level boxes touch (faces, edges, corners),
                 periodic images included and marked as such (brute force;
                 neighbours_by_ring gives the same from an ownership table and
                 serves the plan restatements on large grids)
  refinement     MRAG/MRAGscience/MRAGRefiner_SpaceExtension.h:25-67 and
                 MRAG/MRAGcore/MRAGRefiner.h:43-197
  compression    MRAG/MRAGcore/MRAGCompressor.h:67-146
"""
from __future__ import annotations

import itertools

import numpy as np


def canonical(rows):
    r = np.asarray(rows, np.int32).reshape(-1, 4)
    return np.ascontiguousarray(r[np.lexsort((r[:, 1], r[:, 2], r[:, 3], r[:, 0]))])


def neighbours(rows):
    """list over leaves of sorted [(j, wrapped)]"""
    r = canonical(rows)
    M = int(r[:, 0].max())
    N = 1 << M
    lo = (r[:, 1:].astype(np.int64) << (M - r[:, :1])).astype(np.int64)
    size = (np.int64(1) << (M - r[:, 0].astype(np.int64)))[:, None].repeat(3, 1)
    shifts = np.array(list(itertools.product((-1, 0, 1), repeat=3)), np.int64) * N  # (27, 3)
    out = []
    for a in range(len(r)):
        blo = lo[None, :, :] + shifts[:, None, :]  # (27, n, 3)
        gap = np.where(lo[a] < blo, blo - lo[a] - size[a], lo[a] - blo - size[None, :, :])
        adj = np.all(gap <= 0, axis=2)  # (27, n)
        adj[13, a] = False  # the block itself, unshifted
        s = set()
        for k, b in zip(*np.nonzero(adj)):
            s.add((int(b), bool(k != 13)))
        out.append(sorted(s))
    return out


_RING = {}


def _ring_offsets(size):
    """Cells of the one-cell shell around a cube of `size` cells."""
    if size not in _RING:
        g = np.indices((size + 2,) * 3).reshape(3, -1).T - 1
        inner = np.all((g >= 0) & (g < size), axis=1)
        _RING[size] = g[~inner][:, ::-1]  # (x, y, z) order
    return _RING[size]


def neighbours_by_ring(rows):
    """neighbours() computed from a finest-level ownership table: a leaf
    touches every leaf owning a cell of the one-cell shell around it (a ring
    cell outside the domain is a periodic image).  Same result as the
    brute-force box test, in time linear in the grid size."""
    r = canonical(rows)
    M = int(r[:, 0].max())
    N = 1 << M
    owner = np.full((N, N, N), -1, np.int64)  # [z, y, x]
    lo = (r[:, 1:].astype(np.int64) << (M - r[:, :1])).astype(np.int64)
    size = (1 << (M - r[:, 0])).astype(np.int64)
    for a in range(len(r)):
        x, y, z = lo[a]
        s = size[a]
        owner[z:z + s, y:y + s, x:x + s] = a
    out = []
    for a in range(len(r)):
        c = lo[a][None, :] + _ring_offsets(int(size[a]))
        wrapped = np.any((c < 0) | (c >= N), axis=1)
        c = c % N
        j = owner[c[:, 2], c[:, 1], c[:, 0]]
        pairs = {(int(b), bool(w)) for b, w in zip(j.tolist(), wrapped.tolist()) if not (b == a and not w)}
        out.append(sorted(pairs))
    return out


def refine_victims(rows, flagged, max_level):
    r = canonical(rows)
    nb = neighbours_by_ring(r)
    flag = list(bool(f) for f in flagged)
    lev = r[:, 0]
    for s in range(len(r)):
        if any((not w) and lev[j] > lev[s] for j, w in nb[s]):
            flag[s] = True
    for s in range(len(r)):
        if lev[s] >= max_level:
            flag[s] = False
    victims = {s for s in range(len(r)) if flag[s]}
    front = set(victims)
    while front:
        new = set()
        for v in front:
            for j, _ in nb[v]:
                if abs(lev[j] - (lev[v] + 1)) > 1 and j not in victims:
                    victims.add(j)
                    new.add(j)
        front = new
    return sorted(victims)


def collapse_parents(rows, flagged, min_level=0):
    r = canonical(rows)
    nb = neighbours_by_ring(r)
    index = {tuple(x): i for i, x in enumerate(r.tolist())}
    out = []
    parents = {(L - 1, x >> 1, y >> 1, z >> 1) for L, x, y, z in r.tolist() if L > 0}
    for P in sorted(parents):
        L, px, py, pz = P
        if L < min_level:
            continue
        kids = [index.get((L + 1, 2 * px + (o & 1), 2 * py + ((o >> 1) & 1), 2 * pz + ((o >> 2) & 1))) for o in range(8)]
        if any(k is None or not flagged[k] for k in kids):
            continue
        if all(abs(int(r[j, 0]) - L) <= 1 for k in kids for j, _ in nb[k]):
            out.append(P)
    return sorted(out)


def apply_refine(rows, victims):
    r = canonical(rows).tolist()
    vs = set(victims)
    out = [x for i, x in enumerate(r) if i not in vs]
    for v in victims:
        L, x, y, z = r[v]
        for o in range(8):
            out.append([L + 1, 2 * x + (o & 1), 2 * y + ((o >> 1) & 1), 2 * z + ((o >> 2) & 1)])
    return canonical(out)


def apply_collapse(rows, parents):
    r = canonical(rows).tolist()
    kill = set()
    for L, px, py, pz in parents:
        for o in range(8):
            kill.add((L + 1, 2 * px + (o & 1), 2 * py + ((o >> 1) & 1), 2 * pz + ((o >> 2) & 1)))
    out = [x for x in r if tuple(x) not in kill] + [list(p) for p in parents]
    return canonical(out)


def balanced_random_grid(rng, base_level=1, max_level=3, nrefine=4):
    """A 2:1-balanced grid: random refinements of a uniform grid, each followed
    by refining any leaf that has a neighbour two levels finer."""
    n = 1 << base_level
    rows = canonical([[base_level, x, y, z] for z in range(n) for y in range(n) for x in range(n)])
    for _ in range(nrefine):
        cand = [i for i in range(len(rows)) if rows[i, 0] < max_level]
        if not cand:
            break
        rows = apply_refine(rows, [int(rng.choice(cand))])
        while True:
            nb = neighbours_by_ring(rows)
            bad = sorted({j for s in range(len(rows)) for j, _ in nb[s] if rows[s, 0] - rows[j, 0] > 1})
            if not bad:
                break
            rows = apply_refine(rows, bad)
    return rows


def cell_positions(rows):
    """(nleaves, 8, 8, 8, 3) vertex positions in the unit cube (x = origin + h*i)."""
    r = canonical(rows)
    i = np.arange(8)
    out = np.zeros((len(r), 8, 8, 8, 3))
    for b, (L, x, y, z) in enumerate(r.tolist()):
        h = 2.0 ** -L / 8
        zz, yy, xx = np.meshgrid(i, i, i, indexing="ij")
        out[b, ..., 0] = x * 2.0 ** -L + h * xx
        out[b, ..., 1] = y * 2.0 ** -L + h * yy
        out[b, ..., 2] = z * 2.0 ** -L + h * zz
    return out
