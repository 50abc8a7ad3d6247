"""GPU parity of the batched population solve (gh_batch):

every sample of a mixed batch (three anatomies, grids of 16^3, 24^3 and 32^3
cells, different Dw, rho, Tend and tumour centres, so different step counts)
must equal
  * the single-solve path through gh_ctx (gh_ic_from_anatomy, gh_schedule with
    cfl 0.99, gh_rd_prepare, gh_rd_steps, gh_likelihood) bit for bit, for the
    power-of-two grids that path supports, and
  * the oracle (its _ic restatement, uniform RD steps and likelihood) bit for
    bit on the state, for every grid including the 24^3 one.
"""
from __future__ import annotations

import numpy as np
import pytest

import gliomasolver_amd as gh
from gliomasolver_amd import FIELD_PG, FIELD_PHI, FIELD_PW

pytestmark = pytest.mark.gpu
LIK = (0.000361, 0.85, 0.7, 0.25, 2.0)


def say(*a):
    print("[gpu]", *a, flush=True)


def anatomy(shape, seed):
    """Seeded phantom of twice the largest side, every second voxel, cropped."""
    m = 2 * max(shape)
    gm, wm, csf, _ = gh.phantom(m, seed)
    nz, ny, nx = shape
    return tuple(np.ascontiguousarray(a[::2, ::2, ::2][:nz, :ny, :nx]) for a in (gm, wm, csf))


def observations(n, seed):
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(*(np.arange(n),) * 3, indexing="ij")
    u = np.clip(1.1 - np.sqrt((x - 0.45 * n) ** 2 + (y - 0.5 * n) ** 2 + (z - 0.4 * n) ** 2) / (0.2 * n), 0, 1)
    t1 = (u >= 0.7).astype(np.float64)
    fl = (u >= 0.25).astype(np.float64)
    for m in (t1, fl):
        f = rng.uniform(size=m.shape) < 0.05
        m[f] = 1 - m[f]
    pet = np.where(t1 > 0, np.clip(u + rng.normal(0, 0.05, u.shape), 0, None), 0.0)
    return t1, fl, pet


ANAT = [(4, (36, 40, 44), 21), (3, (30, 24, 28), 22), (2, (16, 16, 16), 23)]   # (bpd, data shape, seed)
SAMPLES = [  # anatomy, Dw, rho, Tend, icx, icy, icz
    (0, 0.05, 0.2, 6.0, 0.45, 0.5, 0.42),
    (1, 0.03, 0.1, 9.0, 0.5, 0.45, 0.5),
    (0, 0.12, 0.3, 3.0, 0.4, 0.55, 0.45),
    (2, 0.02, 0.4, 12.0, 0.5, 0.5, 0.5),
    (1, 0.08, 0.05, 2.0, 0.45, 0.5, 0.45),
    (0, 0.01, 0.25, 10.0, 0.47, 0.47, 0.47),
]


@pytest.fixture(scope="module", params=["f32", "f64"])
def batch_run(request):
    real = request.param
    say("batch", real, "%d samples on 32^3/24^3/16^3" % len(SAMPLES))
    b = gh.Batch(0, real)
    data = []
    for bpd, shape, seed in ANAT:
        gm, wm, csf = anatomy(shape, seed)
        aid = b.add_anatomy(bpd, gm, wm, csf)
        obs = observations(8 * bpd, seed + 100)
        b.set_observations(aid, *obs)
        data.append((bpd, gm, wm, csf, obs))
    anat = [s[0] for s in SAMPLES]
    params = np.array([s[1:] for s in SAMPLES], np.float64)
    b.set_samples(anat, params)
    ms = b.run()
    lik = b.likelihood(LIK, stepPET=2, stepTi=1)
    phis = [b.get_phi(i) for i in range(len(SAMPLES))]
    infos = [b.sample_info(i) for i in range(len(SAMPLES))]
    b.close()
    say("batch device time %.2f ms, steps" % ms, [i[3] for i in infos])
    return real, data, params, phis, lik, infos


def test_batch_matches_single_solve(batch_run):
    real, data, params, phis, lik, infos = batch_run
    for i, (a, *p) in enumerate(SAMPLES):
        bpd, gm, wm, csf, obs = data[a]
        if bpd & (bpd - 1):
            continue                      # the single-solve grid needs a power of two
        with gh.Context(0, real) as ctx:
            ctx.grid_init_uniform(bpd)
            ctx.ic_from_anatomy(gm, wm, csf, tuple(p[3:6]))
            L = float(np.float32(max(gm.shape) * 0.1)) if real == "f32" else max(gm.shape) * 0.1
            Dw, Dg, dt, ns = gh.schedule(real, p[0], L, p[2], 8 * bpd, 0.99)
            assert (Dw, Dg, dt, ns) == infos[i]
            ctx.rd_prepare(Dw, Dg, float(np.float32(p[1])) if real == "f32" else p[1])
            ctx.rd_steps(ns, dt)
            single = ctx.get_field(FIELD_PHI)
            ctx.lik_set_data(*obs)
            tot, comps, _ = ctx.likelihood(LIK, stepPET=2, stepTi=1)
        assert np.array_equal(phis[i], single), "sample %d" % i
        assert lik[i, 0] == tot and tuple(lik[i, 1:]) == comps, "sample %d" % i


def test_batch_matches_oracle(batch_run, oracle):
    real, data, params, phis, lik, infos = batch_run
    t = np.float32 if real == "f32" else np.float64
    for i, (a, *p) in enumerate(SAMPLES):
        bpd, gm, wm, csf, obs = data[a]
        n = 8 * bpd
        pw, pg, _, phi0 = oracle.ic(bpd, gm, wm, csf, tuple(p[3:6]), real)
        Dw, Dg, dt, ns = infos[i]
        ref = oracle.rd_steps(phi0, pw, pg, t(Dw), t(Dg), t(p[1]), t(1.0 / n), dt, ns)
        assert np.array_equal(phis[i], ref), "sample %d (n=%d)" % (i, n)
        tot, comps, _ = oracle.likelihood(ref.astype(np.float64), obs[2], obs[0], obs[1], stepPET=2, stepTi=1)
        assert abs(lik[i, 0] - tot) <= 1e-9 * max(1.0, abs(tot)), "sample %d" % i
    assert len({i[3] for i in infos}) > 3      # the samples really had different step counts


def test_tmcmc_glioma_app(tmp_path):
    """The TMCMC driver end to end on a 16^3 grid: a few generations of a small
    population, each generation one batched solve; generation 0's stored
    log-likelihoods equal single solves of the same points through gh_ctx."""
    import os
    import subprocess

    from conftest import ROOT
    from gliomasolver_amd.datfile import read_dat, write_dat

    gm, wm, csf = anatomy((20, 20, 20), 31)
    for name, a in (("GM", gm), ("WM", wm), ("CSF", csf)):
        write_dat(str(tmp_path / (name + ".dat")), a)
    t1, fl, pet = observations(16, 32)
    for name, a in (("tumT1c", t1), ("tumFLAIR", fl), ("tumPET", pet)):
        write_dat(str(tmp_path / (name + ".dat")), a)
    bounds = [(-3.0, -1.5), (-2.5, -1.0), (2.0, 3.0), (-0.9, -0.6), (-0.8, -0.6), (-0.9, -0.6), (-4.0, -1.4),
              (-0.5, 0.0), (-0.5, -0.2), (-3.0, -0.5), (-3.0, -2.3)]
    par = ["Nth 11", "MaxStages 3", "PopSize 48", "MinChainLength 0", "MaxChainLength 1", "Bdef 0 4"]
    par += ["B%d %r %r" % (i, lo, hi) for i, (lo, hi) in enumerate(bounds)]
    par += ["TolCOV 1", "bbeta 0.2", "seed 280675"]
    (tmp_path / "tmcmc_glioma.par").write_text("\n".join(par) + "\n")
    say("tmcmc_glioma 48 samples x 3 generations on 16^3")
    r = subprocess.run([os.path.join(ROOT, "gliomasolver_amd", "bin", "tmcmc_glioma"), "-PatFileName",
                        str(tmp_path) + "/", "-obs", str(tmp_path), "-bpd", "2", "-stepPET", "1"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    say(r.stdout.strip().splitlines()[-1])
    db0 = np.loadtxt(tmp_path / "curgen_db_000.txt")
    assert db0.shape == (48, 12)
    assert (tmp_path / "curgen_db_002.txt").exists() and (tmp_path / "fitness.txt").exists()
    for i in range(3):
        x = db0[i, :11]
        s, lk = gh.glioma_params(x[None], "f32")
        s, lk = s[0], lk[0]
        with gh.Context(0, "f32") as ctx:
            ctx.grid_init_uniform(2)
            ctx.ic_from_anatomy(gm, wm, csf, tuple(s[3:6]))
            L = float(np.float32(20 * 0.1))
            Dw, Dg, dt, ns = gh.schedule("f32", s[0], L, s[2], 16, 0.99)
            ctx.rd_prepare(Dw, Dg, s[1])
            ctx.rd_steps(ns, dt)
            ctx.lik_set_data(t1, fl, pet)
            nll, _, _ = ctx.likelihood(tuple(lk), stepPET=1, stepTi=1)
        F = gh.glioma_fitness(np.array([nll]))[0]
        assert abs(F - db0[i, 11]) <= 1e-9 * max(1.0, abs(F)), (i, F, db0[i, 11])


@pytest.mark.parametrize("real", ["f32", "f64"])
def test_batch_96_grid_long_run_matches_oracle(real, oracle):
    """The 96^3 grid of the TMCMC stage (a bpd that is not a power of two, the
    stage's most common size) over 240 Euler steps from a Gaussian seed near a
    tissue boundary: the state equals the oracle bit for bit, and the grid the
    batch runs on is the oracle's _ic mapping of the same data in the same
    precision (gh_batch_get_anatomy)."""
    from gliomasolver_amd import workloads as wl

    t = np.float32 if real == "f32" else np.float64
    gm, wm, csf, _ = gh.phantom(96, 5)
    params = (1.3e-5, 0.08, 150.0, 80.0, 51.0, 37.0)
    N = 240
    say("batch 96^3", real, "%d steps" % N)
    b = gh.Batch(0, real)
    aid = b.add_anatomy(12, gm, wm, csf)
    b.set_adjoint_samples([aid], [params], [N])
    b.run()
    phi = b.get_phi(0)
    pw_b, pg_b = b.get_anatomy(aid)
    b.close()
    pw, pg, _, _ = oracle.ic(12, gm, wm, csf, None, real)
    assert np.array_equal(pw_b, pw) and np.array_equal(pg_b, pg)
    u0 = gh.gaussian_ic(real, pw, pg, params[3:], 1.0, 0.1)
    ref = oracle.rd_steps(u0, pw, pg, t(params[0]), t(params[0] / 10.0), t(params[1]), t(1.0 / 96), params[2] / N, N)
    assert np.array_equal(phi, ref)
    assert phi.max() > 0.5                   # the tumour grew
