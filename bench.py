"""Benchmark of the reaction-diffusion hot path on one MI355X.

Workload: uniform 128^3 grid (4,096 blocks of 8^3), FP32 tumour state, seeded
brain phantom, Dw ~ logU[0.002, 0.015] cm^2/day and rho ~ logU[0.002, 0.2]/day
(seeded draw), T = 400 days, explicit Euler with dt = 0.9 h^2/(6 Dw).

One bench "step" is one full forward solve (all Euler steps to T) starting from
the same initial condition, which is restored from a device-side checkpoint.
Inputs are resident in HBM before the timed region starts.

value      = leaf cells x Euler steps x solves / wall seconds (cell-updates/s)
roofline   = 16 B per cell-update (read u, write u, read p_w, read p_g) x cell-
             updates / device time of the step launches (HIP events on the
             context stream), against 8 TB/s; per launch: 16 B x cells x the
             launch's Euler steps (two for the temporally blocked kernel)
adjoint    = forward + likelihood + discrete adjoint (six gradients) of the
             same 128^3 solve, checkpoints every 64 steps: reverse cell-steps/s
             over whole solves, against 52 B per cell-step (forward 16 +
             reverse 20 + tape recompute 16)
batched    = one TMCMC stage (forward + likelihood + adjoint of every sample,
             1,024 samples on 64^3/96^3/128^3 grids) solved by gh_batch in
             memory-sized chunks; evals/s against the same samples one call
             each (timed on the first samples and scaled by their share of
             the stage's cell-steps); FP32 loss/gradient error against the
             FP64 solve of the first samples
cpu_baseline = the reference's ReactionDiffusionOperator + UpdateTumor
             (oracle/_ref/rd_ref_f32, built from the reference sources) over
             MRAG-shaped data -- 27-field AoS Cells, per-block labs loaded from
             ghost instructions -- on the host cores given to one GPU (16), on
             a bounded sample of the same workload; the oracle restatement
             (single thread) when that build is absent.
accuracy   = the FP32 solve against the FP64 solve of the same phantom, IC and
             parameters (each with its own Real schedule): rel-L2 and max-abs
             of phi every 50 days to T = 400 (the FP64 GPU solve equals the
             FP64 CPU oracle bit for bit, tests/test_rd_gpu.py)
uniform256 = the same RD step on a 256^3 grid, whose 400 MB working set
             exceeds the 256 MB Infinity Cache: the HBM-resident rate
adaptive   = the wavelet-adaptive run (levels 3-5 of 8^3 blocks, 256^3
             finest, order-4 wavelets, refine 1e-4 / compress 1e-6 every 10
             steps) over its first epochs: leaf cell-updates/s

Prints exactly one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL_UPDATE_F32 = 16
METRIC = "RD cell-updates/s (% HBM roofline); batched fwd+adjoint evals/s vs plain"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "rd_step2f_uniform128_f32_traffic.json")


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


HOST_SHARE = 16  # host threads a one-GPU job may use on the GPU box


def cpu_threads():
    """The host threads of one GPU's share: at most 16 on the GPU box (it
    exposes more CPUs than one GPU's job may load), fewer if this process is
    pinned to fewer."""
    try:
        visible = len(os.sched_getaffinity(0))
    except AttributeError:
        visible = os.cpu_count() or 1
    return max(1, min(HOST_SHARE, visible))


def cpu_baseline(case, pw, pg, phi0, target_s=12.0):
    """Bounded sample of the same workload on the host CPU."""
    ref = os.path.join(ROOT, "oracle", "_ref", "rd_ref_f32")
    threads = cpu_threads()
    if os.path.exists(ref):
        with tempfile.TemporaryDirectory() as tmp:
            inp, out = os.path.join(tmp, "in.bin"), os.path.join(tmp, "out.bin")

            def run(k):
                with open(inp, "wb") as f:
                    f.write(np.array([case.n], np.int32).tobytes())
                    f.write(np.array([k], np.int64).tobytes())
                    f.write(np.array([case.Dw, case.Dg, case.rho, 1.0 / case.n, case.dt], np.float64).tobytes())
                    for a in (pw, pg, phi0):
                        f.write(np.ascontiguousarray(a, np.float32).tobytes())
                r = subprocess.run([ref, inp, out, str(threads)], capture_output=True, text=True, check=True,
                                   timeout=300)
                return float(r.stdout.split("elapsed_s=")[1].split()[0])

            t2 = run(2)
            k = int(max(2, min(20000, target_s / max(t2 / 2, 1e-6))))
            el = run(k)
            return {"value": case.ncells * k / el, "unit": "cell-updates/s", "cores": threads,
                    "host_cpus_visible": os.cpu_count(), "kind": "reference",
                    "sample": "%d Euler steps of the 128^3 FP32 workload (%.1f s): the reference's "
                              "ReactionDiffusionOperator + UpdateTumor over 8^3 blocks of 27-field Cells, labs "
                              "loaded from ghost instructions, block ranges on %d threads (one GPU's share of "
                              "the %d host CPUs)" % (k, el, threads, os.cpu_count() or 0)}
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from conftest import Oracle  # checker library, used here only as the CPU baseline
    o = Oracle()
    t0 = time.perf_counter()
    o.rd_steps(phi0, pw, pg, case.Dw, case.Dg, case.rho, 1.0 / case.n, case.dt, 2)
    t2 = time.perf_counter() - t0
    k = int(max(2, min(20000, target_s / max(t2 / 2, 1e-6))))
    t0 = time.perf_counter()
    o.rd_steps(phi0, pw, pg, case.Dw, case.Dg, case.rho, 1.0 / case.n, case.dt, k)
    el = time.perf_counter() - t0
    return {"value": case.ncells * k / el, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": "%d Euler steps of the 128^3 FP32 workload (%.1f s), oracle restatement, 1 thread" % (k, el)}


def accuracy_leg(case, marks=(50, 100, 150, 200, 250, 300, 350, 400)):
    """FP32 state against FP64 state of the same phantom, IC and parameters,
    each advanced with its own Real schedule to every mark (days): rel-L2 and
    max-abs of phi.  The FP64 solve runs through the same kernels in FP64,
    which equal the FP64 CPU oracle bit for bit."""
    import gliomasolver_amd as gh
    from gliomasolver_amd import FIELD_PHI
    from gliomasolver_amd import workloads as wl

    c64 = wl.make_case(case.n, "f64", case.seed, Dw_phys=case.Dw_phys, rho=case.rho, tend=case.tend, cfl=case.cfl)
    x32, x64 = wl.load_context(case), wl.load_context(c64)
    done32 = done64 = 0
    rows = []
    for T in marks:
        n32 = gh.schedule("f32", case.Dw_phys, wl.L_CUBE, float(T), case.n, case.cfl)[3]
        n64 = gh.schedule("f64", case.Dw_phys, wl.L_CUBE, float(T), case.n, case.cfl)[3]
        x32.rd_steps(n32 - done32, case.dt)
        x64.rd_steps(n64 - done64, c64.dt)
        done32, done64 = n32, n64
        u32 = x32.get_field(FIELD_PHI).astype(np.float64)
        u64 = x64.get_field(FIELD_PHI)
        d = u32 - u64
        rows.append({"t_days": T, "steps_f32": n32, "steps_f64": n64,
                     "rel_l2": float(np.linalg.norm(d) / np.linalg.norm(u64)), "max_abs": float(np.abs(d).max())})
    x32.close()
    x64.close()
    out = {"what": "phi of the FP32 solve vs the FP64 solve (same phantom, IC, Dw, rho; each precision's own dt and "
                   "step count) every 50 days; the FP64 GPU solve is the FP64 CPU oracle's result bit for bit",
           "bar_rel_l2": 1e-3, "max_rel_l2": max(r["rel_l2"] for r in rows), "marks": rows}
    log("accuracy", json.dumps(out))
    return out


def uniform256_leg(case, steps=1000):
    """The RD step on 256^3 (32,768 blocks): two states (134 MB) and the 16 B
    records (268 MB) exceed the 256 MB Infinity Cache, so each launch streams
    from HBM.  Same parameters as the headline case; a fixed step count."""
    from gliomasolver_amd import workloads as wl

    c = wl.make_case(256, "f32", case.seed, Dw_phys=case.Dw_phys, rho=case.rho, tend=case.tend, cfl=case.cfl)
    ctx = wl.load_context(c)
    ctx.checkpoint_save(0)
    tile = ctx.rd_autotune(c.dt, 20)
    ctx.checkpoint_load(0)
    ctx.rd_steps(100, c.dt)  # warm-up
    ctx.sync()
    ctx.checkpoint_load(0)
    ctx.sync()
    t0 = time.perf_counter()
    kernel_ms = ctx.rd_steps(steps, c.dt, timed=True)
    ctx.sync()
    wall = time.perf_counter() - t0
    ctx.close()
    updates = c.ncells * steps
    achieved = BYTES_PER_CELL_UPDATE_F32 * updates / (kernel_ms * 1e-3) / 1e9
    out = {"workload": "uniform 256^3 RD, FP32, %d Euler steps (working set 400 MB > 256 MB Infinity Cache)" % steps,
           "value": updates / wall, "unit": "cell-updates/s", "schedule": "two Euler steps per launch" if tile == 0
           else "%d blocks per workgroup" % tile,
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4)}}
    log("uniform256", json.dumps(out))
    return out


BATCH_LIK = (0.0025, 1.0, 0.7, 0.25, 2.0)  # PET noise sigma 0.05, PET scale 1; T1Gd/FLAIR thresholds 0.7/0.25


def _stage_memory(cells, nsteps, K, real_size=4):
    """Device bytes of one sample in gh_batch_adjoint: two states, statics (4+1+4 Real),
    two FP64 adjoint states, the (K+1)-state tape and ceil(N/K) checkpoints."""
    per_cell = real_size * (2 + 9 + (K + 1) + -(-nsteps // K)) + 16
    return cells * per_cell


ADJ_BYTES_PER_CELL_STEP_F32 = 52  # forward 16 + reverse 20 + tape recompute 16


def adjoint_leg(ctx, case, final, solves, K):
    """Forward + likelihood + discrete adjoint (dJ/d(Dw, rho, T, x0, y0, z0)) of
    the uniform case from its IC, against observations thresholded from its own
    final state (T1Gd u >= 0.7, FLAIR u >= 0.25, 5 % flips, PET = u + N(0, 0.05^2)
    inside T1Gd).  Rate = reverse cell-steps over the wall clock of whole solves."""
    rng = np.random.default_rng(41)
    u = final.astype(np.float64)
    t1 = (u >= 0.7).astype(np.float64)
    fl = (u >= 0.25).astype(np.float64)
    for m in (t1, fl):
        f = rng.uniform(size=m.shape) < 0.05
        m[f] = 1 - m[f]
    pet = np.where(t1 > 0, u + rng.normal(0, 0.05, u.shape), 0.0)
    ctx.lik_set_data(t1, fl, pet)
    T = case.nsteps * case.dt
    ic = [float(v) for v in case.seed_voxel]
    ctx.checkpoint_load(0)
    ctx.rd_adjoint(case.nsteps, T, ic, 1.0, K)  # warm-up (buffers)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(solves):
        ctx.checkpoint_load(0)
        loss, grad = ctx.rd_adjoint(case.nsteps, T, ic, 1.0, K)
    ctx.sync()
    wall = (time.perf_counter() - t0) / solves
    cell_steps = case.ncells * case.nsteps
    gbs = ADJ_BYTES_PER_CELL_STEP_F32 * cell_steps / wall / 1e9
    out = {"workload": "128^3 FP32 forward + likelihood + discrete adjoint of the same solve (gradient w.r.t. Dw, "
                       "rho, T, x0, y0, z0), FP64 adjoint and gradients, checkpoints every %d steps" % K,
           "euler_steps": case.nsteps, "ms_per_solve": round(wall * 1e3, 3),
           "cell_steps_per_s": cell_steps / wall,
           "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(gbs / HBM_PEAK_GBS, 4),
                        "bytes_per_cell_step": ADJ_BYTES_PER_CELL_STEP_F32,
                        "what": "forward 16 + reverse 20 + recompute 16 B per cell-step, whole-solve wall clock"},
           "loss": float(loss[0]), "grad": [float(g) for g in grad]}
    log("adjoint", json.dumps(out))
    return out


def batched_leg(nsamples, K, budget_gb, seq_samples, acc_samples, chunk_cells=0):
    """One TMCMC stage of forward + likelihood + adjoint over a population of
    64^3 / 96^3 / 128^3 samples (1:2:1): batched on the device (gh_batch), in
    chunks that fit the memory budget, against the same samples solved one at a
    time with the same kernels (a batch of one), and the FP32 results against
    the FP64 solve of the first samples."""
    import gliomasolver_amd as gh
    from gliomasolver_amd import workloads as wl

    st = wl.tmcmc_stage(nsamples)
    cells = st.cells
    log("batched stage: %d samples, sizes %s counts %s, %.3e forward cell-steps" %
        (nsamples, st.sizes, np.bincount(st.anatomy, minlength=len(st.sizes)).tolist(), st.cell_steps()))

    def make_batch(real):
        bt = gh.Batch(0, real)
        ids = [bt.add_anatomy(n // 8, *st.anatomies[a]) for a, n in enumerate(st.sizes)]
        return bt, ids

    b, aids = make_batch("f32")
    # observations: thresholded, noisy copies of a ground-truth solve per anatomy
    rng = np.random.default_rng(77)
    truth_p, truth_n = [], []
    for a, n in enumerate(st.sizes):
        Dw = 0.006 / wl.L_CUBE ** 2
        v = st.wm_voxels[a][len(st.wm_voxels[a]) // 2]
        truth_p.append((Dw, 0.05, 300.0, float(v[0]), float(v[1]), float(v[2])))
        truth_n.append(int(np.ceil(300.0 / (0.9 * (1.0 / n) ** 2 / (6 * Dw)))))
    b.set_adjoint_samples(aids, truth_p, truth_n)
    b.run()
    obs = []
    for a, n in enumerate(st.sizes):
        u = b.get_phi(a).astype(np.float64)
        t1 = (u >= 0.7).astype(np.float64)
        fl = (u >= 0.25).astype(np.float64)
        for m in (t1, fl):
            f = rng.uniform(size=m.shape) < 0.05
            m[f] = 1 - m[f]
        pet = np.where(t1 > 0, u + rng.normal(0, 0.05, u.shape), 0.0)
        obs.append((t1, fl, pet))
        b.set_observations(aids[a], *obs[a])
    # chunks under the memory budget and, when given, a cell budget (a chunk whose
    # per-step working set fits the 256 MB Infinity Cache streams from it, not HBM)
    mem = np.array([_stage_memory(int(c), int(n), K) for c, n in zip(cells, st.nsteps)], np.float64)
    chunks, cur, acc, acc_c = [], [], 0.0, 0
    for i in range(nsamples):
        if cur and (acc + mem[i] > budget_gb * 1e9 or (chunk_cells and acc_c + cells[i] > chunk_cells)):
            chunks.append(cur)
            cur, acc, acc_c = [], 0.0, 0
        cur.append(i)
        acc += mem[i]
        acc_c += int(cells[i])
    chunks.append(cur)
    log("batched: %d chunks, largest %.1f GB" % (len(chunks), max(mem[c].sum() for c in chunks) / 1e9))

    def solve(bt, idx):
        bt.set_adjoint_samples(st.anatomy[idx], st.params[idx], st.nsteps[idx])
        return bt.adjoint(BATCH_LIK, ckpt_every=K)

    solve(b, [0, 1])  # warm-up (code objects, buffers)
    b_loss = np.zeros((nsamples, 4))
    b_grad = np.zeros((nsamples, 6))
    dev_ms = 0.0
    t0 = time.perf_counter()
    for c in chunks:
        loss, grad, ms = solve(b, c)
        b_loss[c], b_grad[c] = loss, grad
        dev_ms += ms
    wall = time.perf_counter() - t0
    log("batched: %d samples in %.2f s (device %.2f s)" % (nsamples, wall, dev_ms / 1e3))
    # sequential: the first samples, one per call, same kernels
    seq = list(range(min(seq_samples, nsamples)))
    identical = True
    t0 = time.perf_counter()
    for i in seq:
        loss, grad, _ = solve(b, [i])
        identical &= bool(np.array_equal(loss[0], b_loss[i]) and np.array_equal(grad[0], b_grad[i]))
    seq_wall = time.perf_counter() - t0
    log("sequential: %d samples in %.2f s (identical to the batch: %s)" % (len(seq), seq_wall, identical))
    anat32 = [b.get_anatomy(aids[a]) for a in range(len(st.sizes))]
    b.close()
    # FP32 state against the FP64 solve of the same samples
    accs = list(range(min(acc_samples, nsamples)))
    b64, ids64 = make_batch("f64")
    for a in range(len(st.sizes)):
        b64.set_observations(ids64[a], *obs[a])
    loss64, grad64, _ = solve(b64, accs)
    # the grids the two precisions run on: _ic places cells and picks voxels in
    # Real, so where 8*bpd is not a power of two FP32 and FP64 map some cells to
    # neighbouring voxels (the reference's own behaviour)
    map_diff = []
    for a in range(len(st.sizes)):
        pw64, pg64 = b64.get_anatomy(ids64[a])
        pw32, pg32 = anat32[a]
        map_diff.append(int(((np.abs(pw32 - pw64) > 1e-6) | (np.abs(pg32 - pg64) > 1e-6)).sum()))
    b64.close()
    rel_loss = np.abs(b_loss[accs, 0] - loss64[:, 0]) / np.abs(loss64[:, 0])
    rel_grad = np.abs(b_grad[accs] - grad64) / np.maximum(np.abs(grad64), 1e-300)
    worst = accs[int(np.argmax(rel_loss))]
    gap = loss_gap(st, worst, obs[st.anatomy[worst]], make_batch)
    gap["anatomy_cells_differing_f32_f64"] = map_diff[st.anatomy[worst]]
    same = [k for k, i in enumerate(accs) if map_diff[st.anatomy[i]] == 0]
    # work of the sequential sample against the whole stage's: evals/s of the plain path
    seq_share = float((cells[seq] * st.nsteps[seq]).sum()) / float((cells * st.nsteps).sum())
    seq_eval_s = len(seq) / seq_wall
    seq_stage_s = seq_wall / seq_share  # the whole stage at the sequential rate per cell-step
    out = {
        "workload": "TMCMC stage: %d samples of (Dw, rho, T, x0, y0, z0), grids 64^3/96^3/128^3 = 1:2:1, "
                    "T ~ U[100, 700] d, forward + likelihood + adjoint, FP32 state, FP64 likelihood and "
                    "gradients, checkpoints every %d steps" % (nsamples, K),
        "evals_per_s": nsamples / wall,
        "stage_wall_s": wall,
        "device_s": dev_ms / 1e3,
        "chunks": len(chunks),
        "forward_cell_steps": st.cell_steps(),
        "sequential_evals_per_s": seq_eval_s,
        "sequential_stage_s_est": seq_stage_s,
        "speedup_vs_sequential": seq_stage_s / wall,
        "roofline": {"bytes_per_cell_step": ADJ_BYTES_PER_CELL_STEP_F32,
                     "stage_bytes": ADJ_BYTES_PER_CELL_STEP_F32 * float(st.cell_steps()),
                     "stage_s_at_peak": ADJ_BYTES_PER_CELL_STEP_F32 * float(st.cell_steps()) / (HBM_PEAK_GBS * 1e9),
                     "batched_frac": ADJ_BYTES_PER_CELL_STEP_F32 * float(st.cell_steps()) / (HBM_PEAK_GBS * 1e9) / wall,
                     "sequential_frac": ADJ_BYTES_PER_CELL_STEP_F32 * float(st.cell_steps()) /
                                        (HBM_PEAK_GBS * 1e9) / seq_stage_s,
                     "max_speedup_over_sequential": seq_stage_s /
                                                    (ADJ_BYTES_PER_CELL_STEP_F32 * float(st.cell_steps()) /
                                                     (HBM_PEAK_GBS * 1e9)),
                     "what": "forward 16 + reverse 20 + recompute 16 B per forward cell-step at 8 TB/s: no batching "
                             "can beat the sequential solves by more than their distance from this bound"},
        "sequential_sample": "first %d samples of the stage, one gh_batch call each (%.1f%% of its cell-steps), "
                             "bit-identical results to the batch: %s" % (len(seq), 100 * seq_share, identical),
        "accuracy_vs_fp64": {"samples": len(accs), "loss_rel_max": float(rel_loss.max()),
                             "loss_rel_median": float(np.median(rel_loss)),
                             "grad_rel_median": float(np.median(rel_grad)), "worst_sample": gap,
                             "anatomy_cells_differing_f32_f64": dict(zip([str(n) for n in st.sizes], map_diff)),
                             "samples_on_identical_grids": len(same),
                             "loss_rel_max_identical_grids": float(rel_loss[same].max()) if same else None,
                             "grad_rel_median_identical_grids": float(np.median(rel_grad[same])) if same else None},
    }
    return out


def loss_gap(st, i, obs, make_batch):
    """Why one sample's FP32 loss differs from its FP64 loss: the likelihood's
    Bernoulli terms jump where u crosses T1uc or T2uc (omega^2 = 1 above uc,
    (u - uc)^2 below, HGG_Likelihood.cpp:136-156), so a voxel whose FP32 and
    FP64 states fall on opposite sides moves the loss by a finite amount.
    Counts those voxels and evaluates the loss of the FP32 state with them
    set to their FP64 values: what is left is the smooth part of the gap."""
    import gliomasolver_amd as gh

    states = []
    for real in ("f32", "f64"):
        bt, ids = make_batch(real)
        bt.set_adjoint_samples([st.anatomy[i]], [st.params[i]], [st.nsteps[i]])
        bt.run()
        states.append(bt.get_phi(0).astype(np.float64))
        bt.close()
    u32, u64 = states
    t1, fl, pet = obs
    uc1, uc2 = BATCH_LIK[2], BATCH_LIK[3]
    cross1 = (u32 > uc1) != (u64 > uc1)
    cross2 = (u32 > uc2) != (u64 > uc2)
    mixed = np.where(cross1 | cross2, u64, u32)
    with gh.Context(0, "f64") as ctx:
        l32 = ctx.likelihood_dense(u32, t1, fl, pet, params=BATCH_LIK)[0]
        l64 = ctx.likelihood_dense(u64, t1, fl, pet, params=BATCH_LIK)[0]
        lmix = ctx.likelihood_dense(mixed, t1, fl, pet, params=BATCH_LIK)[0]
    return {"sample": int(i), "grid": int(st.sizes[st.anatomy[i]]), "loss_f64": l64,
            "loss_rel": abs(l32 - l64) / abs(l64), "voxels_crossing_T1uc": int(cross1.sum()),
            "voxels_crossing_T2uc": int(cross2.sum()), "state_max_abs": float(np.abs(u32 - u64).max()),
            "loss_rel_without_crossings": abs(lmix - l64) / abs(l64)}


def adaptive_cpu_baseline(rows, phi, pw, pg, Dw, Dg, rho, dt, epoch_steps, target_s=8.0):
    """Bounded CPU sample of the adaptive workload on the final grid of the GPU
    run: the oracle's multi-level step (per block and step, labs of phi, p_w and
    p_g with interpolated ghosts -- the Cells a BlockLab loads, MRAGBlockLab.h
    :154-301 -- then the operator) and its FWT details of every block (one
    adaptation pass), on one host thread."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from conftest import Oracle  # checker library, used here only as the CPU baseline
    o = Oracle()
    nb = len(rows)
    t0 = time.perf_counter()
    o.amr_rd_steps(rows, 4, phi, pw, pg, np.float32(Dw), np.float32(Dg), np.float32(rho), dt, 1)
    t1 = time.perf_counter() - t0
    k = int(max(1, min(epoch_steps * 20, (target_s / 2) / max(t1, 1e-6))))
    t0 = time.perf_counter()
    o.amr_rd_steps(rows, 4, phi, pw, pg, np.float32(Dw), np.float32(Dg), np.float32(rho), dt, k)
    el = time.perf_counter() - t0
    t0 = time.perf_counter()
    o.amr_details(rows, 4, phi, list(range(nb)))
    fwt_s = time.perf_counter() - t0
    steps_rate = nb * 512 * k / el
    # one epoch = epoch_steps steps + two detail passes (refine, compress)
    epoch_est = epoch_steps * el / k + 2 * fwt_s
    return {"value": steps_rate, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": "%d multi-level Euler steps (%.1f s) and one FWT-details pass (%.2f s) on the final grid of "
                      "the GPU run (%d blocks, levels %d-%d), oracle restatement of the MRAG lab + operator, 1 thread"
                      % (k, el, fwt_s, nb, int(rows[:, 0].min()), int(rows[:, 0].max())),
            "fwt_details_s": fwt_s, "epoch_s_est": epoch_est,
            "cell_updates_per_s_with_adaptation": nb * 512 * epoch_steps / epoch_est}


def adaptive_leg(tend, epoch_steps, max_epochs, cpu=False):
    """The wavelet-adaptive workload: levels 3-5 of 8^3 blocks (256^3 at the
    finest), 4th-order interpolating wavelets, FP32 state, seeded phantom,
    Dw = 0.013 cm^2/day, rho = 0.025/day, Gaussian IC at a WM voxel; start on
    the uniform finest level and compress (tolerance 1e-6), then every
    epoch_steps Euler steps refine (1e-4, one pass) and compress (1e-6)."""
    import gliomasolver_amd as gh
    from gliomasolver_amd import FIELD_PG, FIELD_PHI, FIELD_PW
    from gliomasolver_amd import workloads as wl

    n = 256
    gm, wm, csf, sv = gh.phantom(n, 3)
    Dw, Dg, dt, nsteps = gh.schedule("f32", 0.013, wl.L_CUBE, tend, n, 0.9)
    ctx = gh.Context(0, "f32")
    ctx.grid_configure(4, 5, 3)
    ctx.grid_init_uniform(n // 8)
    ctx.set_anatomy(gm, wm, csf)
    ctx.set_field(FIELD_PHI, gh.gaussian_ic("f32", ctx.get_field(FIELD_PW), ctx.get_field(FIELD_PG), sv, 1.0, 0.1))
    ctx.rd_prepare(Dw, Dg, 0.025)
    t0 = time.perf_counter()
    ctx.compress(1e-6, -1)
    ctx.sync()
    t_init = time.perf_counter() - t0
    log("adaptive: %d blocks after the initial compression (%.2f s)" % (ctx.nblocks(), t_init))
    ctx.adapt_profile(reset=True)
    leaf_updates, steps_done, step_s, epoch_s, blocks = 0, 0, 0.0, 0.0, []
    epochs = min(max_epochs, -(-nsteps // epoch_steps))
    t0 = time.perf_counter()
    for e in range(epochs):
        k = min(epoch_steps, nsteps - steps_done)
        nb = ctx.nblocks()
        blocks.append(nb)
        t1 = time.perf_counter()
        ctx.rd_steps(k, dt)
        ctx.sync()
        t2 = time.perf_counter()
        ctx.refine(1e-4, 1)
        ctx.compress(1e-6, -1)
        ctx.sync()
        t3 = time.perf_counter()
        step_s += t2 - t1
        epoch_s += t3 - t2
        leaf_updates += nb * 512 * k
        steps_done += k
        if (e + 1) % 25 == 0 or e + 1 == epochs:
            log("adaptive: epoch %d/%d, %d blocks, steps %.2f s, epochs %.2f s" % (e + 1, epochs, ctx.nblocks(),
                                                                                   step_s, epoch_s))
    wall = time.perf_counter() - t0
    prof = {k: round(v, 1) for k, v in ctx.adapt_profile().items()}
    log("adaptive: epoch profile (ms)", json.dumps(prof))
    cpu_b = None
    if cpu:
        lev, idx = ctx.blocks()
        rows = np.c_[lev, idx].astype(np.int32)
        fields = [ctx.get_field_blocks(f).reshape(len(rows), 512) for f in (FIELD_PHI, FIELD_PW, FIELD_PG)]
        log("adaptive: cpu baseline ...")
        cpu_b = adaptive_cpu_baseline(rows, *fields, Dw, Dg, 0.025, dt, epoch_steps)
        log("adaptive: cpu baseline", json.dumps(cpu_b))
    ctx.close()
    return {"metric": "adaptive RD leaf cell-updates/s", "value": leaf_updates / wall, "unit": "cell-updates/s",
            "config": {"workload": "wavelet-adaptive levels 3-5 (256^3 finest), order-4 interpolating wavelets, "
                                   "FP32, refine 1e-4 / compress 1e-6 every %d steps" % epoch_steps,
                       "steps": steps_done, "of_total_steps": nsteps, "dt_days": dt, "epochs": epochs,
                       "blocks_min_mean_max": [min(blocks), float(np.mean(blocks)), max(blocks)]},
            "step_only_cell_updates_per_s": leaf_updates / step_s, "step_s": step_s, "epoch_s": epoch_s,
            "epoch_ms_mean": 1e3 * epoch_s / max(1, epochs), "initial_compress_s": t_init, "wall_s": wall,
            "epoch_profile_ms": prof, "cpu_baseline": cpu_b}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed full solves")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched TMCMC-stage leg")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the single-solve adjoint leg")
    ap.add_argument("--adjoint-ckpt", type=int, default=64, help="checkpoint interval of the single-solve adjoint")
    ap.add_argument("--adjoint-solves", type=int, default=3, help="timed adjoint solves")
    ap.add_argument("--batch-samples", type=int, default=1024)
    ap.add_argument("--batch-ckpt", type=int, default=32, help="checkpoint interval of the batched adjoint")
    ap.add_argument("--batch-budget-gb", type=float, default=120.0, help="device memory per batched chunk")
    ap.add_argument("--batch-cells", type=float, default=0, help="cells per batched chunk (0: memory budget only)")
    ap.add_argument("--batch-seq", type=int, default=64, help="samples timed one at a time")
    ap.add_argument("--batch-acc", type=int, default=24, help="samples re-solved in FP64")
    ap.add_argument("--workload", default="uniform128", choices=["uniform128", "adaptive"],
                    help="adaptive: the levels 3-5 wavelet-adaptive run instead (its own JSON line)")
    ap.add_argument("--adaptive-epochs", type=int, default=200)
    ap.add_argument("--no-adaptive", action="store_true", help="skip the adaptive leg of the default line")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the FP32-vs-FP64 accuracy report")
    ap.add_argument("--no-256", action="store_true", help="skip the 256^3 (HBM-resident) uniform leg")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (samples shard with no "
              "intra-stage traffic; see the design notes). Use --gpus 1.", file=sys.stderr)
        return 2

    if a.workload == "adaptive":
        print(json.dumps(adaptive_leg(500.0, 10, a.adaptive_epochs, cpu=not a.no_cpu)), flush=True)
        return 0

    import gliomasolver_amd as gh
    from gliomasolver_amd import FIELD_PG, FIELD_PHI, FIELD_PW
    from gliomasolver_amd import workloads as wl

    case = wl.config_uniform128(a.seed)
    log("workload", json.dumps(case.describe()))
    ctx = wl.load_context(case)
    pw, pg, phi0 = ctx.get_field(FIELD_PW), ctx.get_field(FIELD_PG), ctx.get_field(FIELD_PHI)
    ctx.checkpoint_save(0)
    tile = ctx.rd_autotune(case.dt, 100)
    clamped = ctx.rd_schedule_info()[1]
    log("autotuned schedule: %s" % ("two Euler steps per launch" if tile == 0 else
                                    "%d blocks of 8^3 per workgroup" % tile))
    for i in range(a.warmup):
        ctx.checkpoint_load(0)
        ctx.rd_steps(case.nsteps, case.dt)
    ctx.sync()
    log("warmup done")
    kernel_ms = 0.0
    ctx.sync()
    t0 = time.perf_counter()
    for i in range(a.steps):
        ctx.checkpoint_load(0)
        kernel_ms += ctx.rd_steps(case.nsteps, case.dt, timed=True)
    ctx.sync()
    wall = time.perf_counter() - t0
    log("timed %d solves x %d Euler steps in %.3f s" % (a.steps, case.nsteps, wall))
    final = ctx.get_field(FIELD_PHI)
    adj = None if a.no_adjoint else adjoint_leg(ctx, case, final, a.adjoint_solves, a.adjoint_ckpt)
    ctx.close()

    updates = case.ncells * case.nsteps * a.steps
    value = updates / wall
    steps_per_launch = 2 if tile == 0 else 1
    launches = (case.nsteps // steps_per_launch + case.nsteps % steps_per_launch) * a.steps
    avg_launch_s = kernel_ms * 1e-3 / launches
    # algorithmic bytes of all cell-updates over the device time of all launches
    achieved = BYTES_PER_CELL_UPDATE_F32 * updates / (kernel_ms * 1e-3) / 1e9
    kbase = ("rd_step2f_kernel" if clamped else "rd_step2_kernel") if tile == 0 else "rd_step_kernel"
    traffic, traffic_src = None, None
    if os.path.exists(TRAFFIC_FILE):
        tf = json.load(open(TRAFFIC_FILE))
        if tf.get("kernel") == kbase:
            traffic = tf.get("bytes_per_launch")
            traffic_src = ("stored rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs of this kernel on this workload "
                           "(%s, copy-kernel calibrated; per-launch time there %.3f us)" %
                           (os.path.relpath(TRAFFIC_FILE, ROOT), tf.get("avg_launch_us", float("nan"))))
    kname = ("%s<float> (2 Euler steps per launch)" % kbase if tile == 0 else
             "rd_step_kernel<float,true,%d>" % tile)
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": traffic_src,
            "kernel": kname, "avg_launch_us": round(avg_launch_s * 1e6, 3), "steps_per_launch": steps_per_launch,
            "algorithmic_bytes_per_launch": BYTES_PER_CELL_UPDATE_F32 * case.ncells * steps_per_launch}
    cpu = None
    if not a.no_cpu:
        log("cpu baseline ...")
        cpu = cpu_baseline(case, pw, pg, phi0)
        log("cpu baseline", json.dumps(cpu))
    line = {"metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": a.steps,
            "warmup": a.warmup, "ms_per_step": wall * 1e3 / a.steps, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "uniform 128^3 RD to T=400 d, FP32 state, bit-exact to the reference operator",
                       "grid": "128^3 (4096 blocks of 8^3)", "euler_steps_per_solve": case.nsteps,
                       "Dw_phys": case.Dw_phys, "rho": case.rho, "dt_days": case.dt, "seed": case.seed,
                       "final_phi_sum": float(final.astype(np.float64).sum())},
            "roofline": roof, "cpu_baseline": cpu}
    if cpu:
        line["gpu_over_cpu"] = round(value / cpu["value"], 1)
    if not a.no_accuracy:
        log("accuracy: FP32 vs FP64 every 50 days ...")
        line["accuracy"] = accuracy_leg(case)
    if not a.no_256:
        log("uniform 256^3 ...")
        line["uniform256"] = uniform256_leg(case)
    if adj:
        line["adjoint"] = adj
    if not a.no_adaptive:
        log("adaptive run ...")
        line["adaptive"] = adaptive_leg(500.0, 10, a.adaptive_epochs, cpu=not a.no_cpu)
    if not a.no_batch:
        log("batched TMCMC stage ...")
        line["batched"] = batched_leg(a.batch_samples, a.batch_ckpt, a.batch_budget_gb, a.batch_seq, a.batch_acc,
                                      int(a.batch_cells))
        log("batched", json.dumps(line["batched"]))
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
