"""Benchmark: ABC particles/sec for the plain path (SSA + statistics + errors).

Workload: all five models, the cell design of the SSA workload (2,500 cells
per particle, theta ~ U[0,1), 21 h cycle, 5 age bins, 4 pulses + 5 chases,
Beta(2,8) capture, config priors, Philox seed 0); every particle's statistics
(the 90-value mean/Fano superset) are scored against 2,000 synthetic genes
whose data statistics come from ground-truth particles of the same priors with
SE = |stat| * U[0.05, 0.3].  One step = one batch of P particles per model
through prior -> rates -> SSA -> statistics -> FP64 errors, all in HBM.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--particles P]

Prints one JSON line.  Only rank 0 / one GPU is supported.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ABC particles/sec (SSA+stats+errors); accepted/sec surrogate vs plain GPU"
HBM_PEAK = 8000.0       # GB/s


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--particles", type=int, default=4096, help="particles per model per step")
    ap.add_argument("--genes", type=int, default=2000)
    ap.add_argument("--cells", type=int, default=2500)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the occupancy-default SSA grid")
    ap.add_argument("--no-graph", action="store_true", help="skip the HIP-graph replay of the steps")
    ap.add_argument("--sur-train", type=int, default=48000, help="training particles per model")
    ap.add_argument("--sur-eval", type=int, default=2000, help="held-out particles per model")
    ap.add_argument("--sur-steps", type=int, default=6000, help="Adam steps (batch 8192)")
    ap.add_argument("--no-moments", action="store_true",
                    help="skip the leg of the reference's own algorithm (moment model)")
    ap.add_argument("--mom-particles", type=int, default=32768,
                    help="particles per model of the moment-model leg")
    ap.add_argument("--mom-method", type=int, default=1, choices=[0, 1],
                    help="moment-model integration: 1 = gene block in closed form, "
                         "0 = Dormand-Prince on all nine moments")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run three device copies of 256 MiB (profiler byte-counter scale)")
    return ap.parse_args()


def copy_calibration():
    import torch
    n = 256 * 1024 * 1024 // 4
    src = torch.ones(n, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read_per_launch": 4 * n, "copy_bytes_written_per_launch": 4 * n}))


def main():
    a = parse()
    if a.copy_calibration:
        copy_calibration()
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if a.gpus != 1 or world != 1:
        print("bench.py measures one GPU; runs on several GPUs are not supported "
              "(particles would shard by contiguous ranges of the global particle index, with no "
              "communication: each range's streams are keyed by that index)",
              file=sys.stderr)
        sys.exit(2)
    import torch
    from abc_inference_transcription_amd import _lib
    from abc_inference_transcription_amd.design import config_design
    from abc_inference_transcription_amd.engine import Engine

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    design = config_design(cells=a.cells)
    eng = Engine(design, device=0)
    L = eng.L
    # one explicit (non-null, default-priority) stream shared by torch and libabc, so the
    # timing events below bracket exactly the library's kernels
    stream = torch.cuda.Stream(device=dev)
    torch.cuda.set_stream(stream)
    _lib.check(L.abc_set_stream(eng.ctx, C.c_void_p(stream.cuda_stream)), eng.ctx, "set_stream")
    nb, J = eng.nb, eng.J
    KS = 2 * nb * J
    parts = np.array([nb * J, nb * J], np.int32)
    models = [1, 2, 3, 4, 5]
    P = a.particles

    def ptr(t):
        return C.c_void_p(t.data_ptr())

    # ---- synthetic genes: ground-truth particles of the same priors (untimed)
    G = a.genes
    gt_first = 3_000_000_000  # stream range disjoint from the benchmark particles
    per = (G + len(models) - 1) // len(models)
    data_rows = []
    for m in models:
        npar = eng.n_params(m)
        th = torch.empty((npar, per), dtype=torch.float64, device=dev)
        st = torch.empty((4 * nb + 3 * J, per), dtype=torch.float64, device=dev)
        sup = torch.empty((KS, per), dtype=torch.float64, device=dev)
        _lib.check(L.abc_sample_prior_dev(eng.ctx, m, 0, gt_first + m * per, per, ptr(th)), eng.ctx, "prior")
        _lib.check(L.abc_simulate_dev(eng.ctx, m, 0, gt_first + m * per, per, ptr(th), ptr(st), None,
                                      ptr(sup), None), eng.ctx, "simulate")
        data_rows.append(sup.t().cpu().numpy())
    torch.cuda.synchronize()
    data = np.vstack(data_rows)[:G]
    rng = np.random.default_rng(0)
    se = np.abs(data) * rng.uniform(0.05, 0.3, data.shape)
    data_d = torch.tensor(np.asfortranarray(data).ravel(order="F"), dtype=torch.float64, device=dev)
    se_d = torch.tensor(np.asfortranarray(se).ravel(order="F"), dtype=torch.float64, device=dev)
    parts_c = np.ascontiguousarray(parts)

    # ---- device buffers for one step
    bufs = {}
    for m in models:
        npar = eng.n_params(m)
        bufs[m] = dict(
            theta=torch.empty(npar * P, dtype=torch.float64, device=dev),
            stats=torch.empty((4 * nb + 3 * J) * P, dtype=torch.float64, device=dev),
            sup=torch.empty(KS * P, dtype=torch.float64, device=dev),
            steps=torch.zeros(P, dtype=torch.int64, device=dev),
            err=torch.empty(P * G, dtype=torch.float64, device=dev),
        )
    steps_acc = torch.zeros((), dtype=torch.int64, device=dev)
    # all five models' particles go through one burn-in and one SSA launch per step
    nm = len(models)
    VA = C.c_void_p * nm
    m_models = (C.c_int32 * nm)(*models)
    m_n = (C.c_int64 * nm)(*([P] * nm))
    m_theta = VA(*[bufs[m]["theta"].data_ptr() for m in models])
    m_stats = VA(*[bufs[m]["stats"].data_ptr() for m in models])
    m_sup = VA(*[bufs[m]["sup"].data_ptr() for m in models])
    m_steps = VA(*[bufs[m]["steps"].data_ptr() for m in models])

    def step(it, timed):
        first = it * P
        for m in models:
            _lib.check(L.abc_sample_prior_dev(eng.ctx, m, 0, first, P, ptr(bufs[m]["theta"])),
                       eng.ctx, "prior")
        _lib.check(L.abc_simulate_multi_dev(eng.ctx, nm, m_models, 0, (C.c_int64 * nm)(*([first] * nm)),
                                            m_n, m_theta, m_stats, None, m_sup, m_steps),
                   eng.ctx, "simulate")
        for m in models:
            b = bufs[m]
            _lib.check(L.abc_errors_dev(eng.ctx, P, KS, ptr(b["sup"]), G, ptr(data_d), ptr(se_d),
                                        parts_c.ctypes.data_as(C.c_void_p), 2, 10.0, ptr(b["err"])),
                       eng.ctx, "errors")
            if timed:
                steps_acc.add_(b["steps"].sum())  # same stream: no host sync

    def ssa_time(reset):
        ms, bst, nl = (C.c_double * 2)(), C.c_int64(), C.c_int64()
        _lib.check(L.abc_ssa_time(eng.ctx, reset, ms, C.byref(bst), C.byref(nl)), eng.ctx,
                   "ssa_time")
        return ms[0], ms[1], bst.value, nl.value

    tuned = None
    if not a.no_autotune:
        # grid size of the persistent SSA kernel, tuned on one full step's worth of
        # other particles: a smaller sample is dominated by its longest work items,
        # which run fastest at low occupancy, and would mislead the choice
        blocks, tms = eng.autotune(models=models, n=P, first=50_000_000)
        tuned = {"ssa_blocks": blocks, "sample_ssa_ms": round(tms, 3)}
        print(f"[bench] autotune: {blocks} workgroups", file=sys.stderr, flush=True)
    for w in range(a.warmup):
        step(10_000 + w, False)
    torch.cuda.synchronize()
    ssa_time(1)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    for it in range(a.steps):
        step(it, True)
        print(f"[bench] step {it} enqueued", file=sys.stderr, flush=True)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = ev0.elapsed_time(ev1)
    n_particles = a.steps * P * len(models)
    value = n_particles / (gpu_ms / 1e3)
    burn_ms_total, ssa_ms_total, burn_steps, launches = ssa_time(1)
    total_steps = int(steps_acc.item())
    window_steps = total_steps - burn_steps

    graph = None
    if not a.no_graph:
        # the first steps again, each captured into one HIP graph (updated in
        # place for the next step's particles) and replayed
        g = None
        gsteps = min(a.steps, 2)
        ev2, ev3 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        ev2.record(stream)
        for it in range(gsteps):
            g = eng.capture(lambda: step(it, False), graph=g)
            eng.replay(g)
        ev3.record(stream)
        torch.cuda.synchronize()
        gms = ev2.elapsed_time(ev3)
        eng.drop_graph(g)
        ssa_time(1)
        graph = {"value": round(gsteps * P * len(models) / (gms / 1e3), 3),
                 "ms_per_step": round(gms / gsteps, 3), "steps": gsteps}
        print(f"[bench] graph replay: {graph}", file=sys.stderr, flush=True)

    # ---- roofline of the dominant kernel (ssa_kernel, the labelling window):
    # algorithmic VALU lane-ops of its event steps over its own launch time
    from abc_inference_transcription_amd.roofline import SSA_OPS_PER_STEP, VALU_PEAK_TOPS
    ssa_s = ssa_ms_total / 1e3
    steps_per_s = window_steps / ssa_s
    achieved = steps_per_s * SSA_OPS_PER_STEP / 1e12
    roof = {"kernel": "ssa_kernel", "bound": "valu", "achieved": round(achieved, 3),
            "peak": VALU_PEAK_TOPS, "unit": "Top/s", "frac": round(achieved / VALU_PEAK_TOPS, 4),
            "traffic": None, "steps_per_s": steps_per_s, "ops_per_step": SSA_OPS_PER_STEP,
            "window_steps": window_steps,
            "avg_launch_ms": ssa_ms_total / max(launches, 1), "launches": launches,
            "ssa_share_of_step": ssa_s / (gpu_ms / 1e3),
            "burn_kernel": {"avg_launch_ms": burn_ms_total / max(launches, 1),
                            "steps": burn_steps,
                            "share_of_step": burn_ms_total / gpu_ms}}

    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "particles/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(gpu_ms / a.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (prior draws, Philox seed 0; genes from ground-truth prior particles)",
        "config": {"workload": "plain ABC: 5 models x %d particles/step, 2500 cells, SSA+stats+FP64 errors vs %d genes" % (P, G),
                   "models": models, "particles_per_model_per_step": P, "cells_per_particle": a.cells,
                   "genes": G, "stats": KS, "cycle_h": 21.0, "capture": "Beta(2,8)",
                   "mean_steps_per_particle": total_steps / n_particles},
        "roofline": roof,
        "wall_s": round(wall, 3),
        "ssa_launch": {"workgroups": eng.ssa_blocks(), "wg_lanes": 256, "autotune": tuned},
        "graph_replay": graph,
    }

    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(design, models, P, total_steps / n_particles, a.cpu_seconds)
    if not a.no_surrogate:
        out["surrogate"] = surrogate_leg(eng, stream, a)
    if not a.no_moments:
        out["reference_moment_path"] = moment_leg(stream, a)
    out["model_probs"] = model_probs_leg(eng, stream, a)
    out["data_summary_stats"] = data_stats_leg(eng, stream, a)
    print(json.dumps(out), flush=True)
    eng.close()


def surrogate_leg(eng, stream, a):
    """Accepted pairs/s of the surrogate-screened path against the exhaustive
    plain path on the same held-out particles (top-k per gene), with recall."""
    from abc_inference_transcription_amd import screened_abc as S
    from abc_inference_transcription_amd.surrogate import Surrogate
    sur = Surrogate(2 * eng.nb * eng.J, seed=0, stream=stream)
    res = S.run(eng, sur, G=a.genes, n_train=a.sur_train, n_eval=a.sur_eval, k=0,
                accept_rate=4e-5, z=[4.0, 3.0, 2.5, 2.0], train_steps=a.sur_steps, batch=8192,
                log=lambda *x: print(*x, file=sys.stderr, flush=True))
    sur.close()
    keep = ("plain_accepted_per_s", "screened_accepted_per_s", "speedup_accepted_per_s",
            "recall", "ssa_fraction_avoided", "ssa_work_fraction_avoided", "plain_pairs",
            "eval_particles", "train_particles", "heldout_nll", "heldout_nll_of_mean_predictor",
            "heldout_coverage_2sigma", "heldout_z_rms", "heldout_rmse_std_units_median",
            "train_loss_last", "train_s", "genes", "z", "eps", "accept_rate_target", "by_z")
    return {k: res[k] for k in keep}


def moment_leg(stream, a):
    """The reference's own algorithm (scripts/model.jl moment ODE + the abc_sim
    statistics of abc_simulation.jl, reference design: 20 h cycle, 5 ages at
    the bin centres, 11 conditions, capture downsampling) on the GPU
    (abc_moment_sim_dev) and the oracle's C restatement of it on the host
    cores, both integrating with the same steps (rtol 1e-10; method 1 solves the
    stiff gene block in closed form and takes Dormand-Prince steps over the
    other five moments, method 0 takes them over all nine)."""
    import torch
    from abc_inference_transcription_amd import _lib
    from abc_inference_transcription_amd.design import REF_CHASE, REF_PULSE, reference_design
    from abc_inference_transcription_amd.engine import Engine, MomentCfg
    from abc_inference_transcription_amd.roofline import (FP64_PEAK_TFLOPS,
                                                          MOMENT_FLOPS_PER_STEP_BY_METHOD)
    flops_per_step = MOMENT_FLOPS_PER_STEP_BY_METHOD[a.mom_method]
    dev = torch.device("cuda:0")
    # a context of the reference design, so that the particles are draws of the
    # reference prior (fix_params ranges), as on the CPU side
    eng = Engine(reference_design(cells=55), device=0)
    eng.bind_stream(stream)
    L = eng.L
    ages = [(2 * k + 1) / 10 * 20 for k in range(5)]
    is_chase = [0] * 6 + [1] * 5
    cfg = MomentCfg(ages, REF_PULSE, REF_CHASE, is_chase=is_chase, cycle=20.0,
                    iv=[0.0, 0.5] + [0.0] * 7, method=a.mom_method)
    w = np.ascontiguousarray(np.full(55, 0.2))
    # Beta(2,8) capture in every (age, pulse/chase) group: E b, E b^2, Var b
    cap = np.ascontiguousarray(np.tile([0.2, 6.0 / 110.0, 16.0 / 1100.0], 10))
    n = a.mom_particles
    models = [1, 2, 3, 4, 5]
    bufs = []
    for m in models:
        th = torch.empty(n * eng.n_params(m), dtype=torch.float64, device=dev)
        _lib.check(L.abc_sample_prior_dev(eng.ctx, m, 0, 0, n, C.c_void_p(th.data_ptr())),
                   eng.ctx, "prior")
        bufs.append((m, th, torch.empty(n * 53, dtype=torch.float64, device=dev),
                     torch.zeros(n, dtype=torch.int64, device=dev)))
    wp, cp = w.ctypes.data_as(C.c_void_p), cap.ctypes.data_as(C.c_void_p)

    def run(m, th, st, steps):
        _lib.check(L.abc_moment_sim_dev(eng.ctx, m, n, C.c_void_p(th.data_ptr()), n,
                                        C.byref(cfg.s), cp, wp, 5, 6, C.c_void_p(st.data_ptr()),
                                        None, C.c_void_p(steps.data_ptr())), eng.ctx,
                   "moment_sim")
    run(*bufs[0])  # warm-up (code load, buffer growth)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for b in bufs:
        run(*b)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    steps = int(sum(int(b[3].sum()) for b in bufs))
    value = len(models) * n / (ms / 1e3)
    tflops = steps * flops_per_step / (ms / 1e3) / 1e12
    eng.close()
    print(f"[bench] moment path: {value:.1f} particles/s", file=sys.stderr, flush=True)
    res = {"value": round(value, 3), "unit": "particles/s", "particles": len(models) * n,
           "ms": round(ms, 3), "rtol": 1e-10, "method": a.mom_method,
           "rk_steps_per_particle": steps / (len(models) * n),
           "roofline": {"bound": "fp64 valu", "achieved": round(tflops, 3),
                        "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                        "frac": round(tflops / FP64_PEAK_TFLOPS, 4),
                        "flops_per_step": flops_per_step}}
    if not a.no_cpu:
        res["cpu"] = moment_cpu(models, cfg, w, is_chase, cap, a.cpu_seconds, a.mom_method)
        res["gpu_vs_cpu"] = round(value / res["cpu"]["value"], 1)
    return res


def model_probs_leg(eng, stream, a):
    """get_model_probs for 2,000 genes (scripts/model_probs.jl): two model
    groups with 0..20,000 accepted particles each (synthetic counts, seed 0),
    100 resamples per gene, on the GPU; the oracle's scalar replay of the same
    resamples timed on a sample of genes.  One unit = one resampled label
    (a 64-bit uniform, a multiply-high and the group search)."""
    import torch
    counts = np.random.default_rng(0).integers(0, 20001, size=(2000, 2))
    draws = int(counts.sum()) * 100
    eng.model_probs(counts[:8])  # warm-up
    torch.cuda.synchronize()
    t = time.perf_counter()
    eng.model_probs(counts)
    gpu_s = time.perf_counter() - t
    res = {"genes": 2000, "groups": 2, "resamples": 100, "draws": draws,
           "gpu_ms": round(gpu_s * 1e3, 3), "gpu_draws_per_s": draws / gpu_s,
           "note": "host-array call: includes the copies and the synchronisation"}
    if not a.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        sample = counts[:20]
        t = time.perf_counter()
        O.model_probs(sample)
        cpu_s = time.perf_counter() - t
        cd = int(sample.sum()) * 100
        res["cpu"] = {"draws_per_s": cd / cpu_s, "cores": 1, "kind": "port",
                      "sample": f"20 genes, {cd} draws in {cpu_s:.2f}s"}
        res["gpu_vs_cpu"] = round((draws / gpu_s) / (cd / cpu_s), 1)
    return res


DATA_OPS_PER_DRAW = 50  # half a Philox4x32-10 block (30), a multiply-high, the draw count, the cell's sums


def data_stats_leg(eng, stream, a):
    """get_summary_stats with bootstrap SEs (scripts/data_summary_statistics.jl)
    at the reference data's size: 3,419 genes, 5,422 cells, 5 age clusters,
    11 conditions, 100 resamples; synthetic negative-binomial counts (the
    raw data are not shipped), counts already in device memory.  One unit =
    one resampled cell (each resample draws the pulse cells, the chase cells
    and all cells twice).  The oracle's threaded C restatement runs on the
    first 256 genes."""
    import torch
    from abc_inference_transcription_amd.roofline import VALU_PEAK_TOPS
    rng = np.random.default_rng(0)
    G, n, nc, J, B = 3419, 5422, 5, 11, 100
    age = rng.integers(0, nc, n).astype(np.int32)
    cond = rng.integers(-1, J, n).astype(np.int32)
    u = rng.negative_binomial(2, rng.uniform(0.05, 0.8, G), (n, G)).astype(np.uint32)
    l = rng.negative_binomial(1, rng.uniform(0.2, 0.95, G), (n, G)).astype(np.uint32)
    pulse = np.nonzero((cond >= 0) & (cond <= 5))[0].astype(np.int32)
    chase = np.nonzero(cond >= 6)[0].astype(np.int32)
    w = np.asfortranarray([[np.mean(age[cond == j] == c) for j in range(J)] for c in range(nc)])
    dev = torch.device("cuda:0")
    ut = torch.from_numpy(np.ascontiguousarray(u.T).view(np.int32)).to(dev)
    lt = torch.from_numpy(np.ascontiguousarray(l.T).view(np.int32)).to(dev)
    out = torch.empty(G * (8 * nc + 6 * J), dtype=torch.float64, device=dev)
    vp = lambda x: C.c_void_p(x.ctypes.data)

    def run(ng):
        from abc_inference_transcription_amd import _lib
        _lib.check(eng.L.abc_data_stats_dev(eng.ctx, ng, n, C.c_void_p(ut.data_ptr()),
                                            C.c_void_p(lt.data_ptr()), nc, vp(age), J, vp(cond),
                                            vp(pulse), len(pulse), vp(chase), len(chase), vp(w),
                                            B, 0, C.c_void_p(out.data_ptr())), eng.ctx,
                   "data_stats")
    run(8)  # warm-up
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    run(G)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    draws = G * B * (len(pulse) + len(chase) + 2 * n)
    tops = draws * DATA_OPS_PER_DRAW / (ms / 1e3) / 1e12
    res = {"genes": G, "cells": n, "resamples": B, "value": round(G / (ms / 1e3), 1),
           "unit": "genes/s", "ms": round(ms, 3), "resampled_cells": draws,
           "roofline": {"bound": "valu", "achieved": round(tops, 3), "peak": VALU_PEAK_TOPS,
                        "unit": "Top/s", "frac": round(tops / VALU_PEAK_TOPS, 4),
                        "ops_per_resampled_cell": DATA_OPS_PER_DRAW}}
    print(f"[bench] data summary statistics: {res['value']:.1f} genes/s", file=sys.stderr,
          flush=True)
    if not a.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        threads = min(os.cpu_count() or 1, 16)
        k = 256
        t = time.perf_counter()
        O.data_stats(u[:, :k], l[:, :k], age, cond, nc, J, pulse, chase, w, n_boot=B,
                     n_threads=threads)
        cpu_s = time.perf_counter() - t
        res["cpu"] = {"value": round(k / cpu_s, 3), "unit": "genes/s", "cores": threads,
                      "kind": "port", "sample": f"{k} genes in {cpu_s:.2f}s on {threads} threads"}
        res["gpu_vs_cpu"] = round(res["value"] / res["cpu"]["value"], 1)
    return res


def moment_cpu(models, cfg, w, is_chase, cap, budget_s, method):
    """The oracle's threaded C restatement of the same moment abc_sim on the
    host cores, over a bounded sample: 16 particles per model, round-robin,
    from the start of each model's prior stream."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from abc_inference_transcription_amd.design import reference_design
    threads = min(os.cpu_count() or 1, 16)
    od = O.OracleDesign(reference_design(cells=55))
    ocfg = O.moment_cfg(cfg.ages, cfg.pulse, cfg.chase, cycle=20.0, iv=[0.0, 0.5] + [0.0] * 7,
                        method=method)
    done, t_used, it = 0, 0.0, 0
    while t_used < budget_s and it < 500:
        m = models[it % len(models)]
        th = O.sample_prior(od, m, 0, (it // len(models)) * 16, 16)
        t = time.perf_counter()
        O.moment_sim(m, th, ocfg, w.reshape(5, 11, order="F"), is_chase=is_chase, capture=cap,
                     n_threads=threads)
        t_used += time.perf_counter() - t
        done += 16
        it += 1
    return {"value": round(done / t_used, 3), "unit": "particles/s", "cores": threads,
            "kind": "port", "sample": f"{done} particles (16 per model round-robin) in "
                                      f"{t_used:.1f}s on {threads} threads"}


def cpu_baseline(design, models, P, gpu_steps_per_particle, budget_s):
    """The oracle's scalar replay of the same SSA (the identical Philox steps)
    on the host cores, over a bounded sample of the benchmark's particles:
    the same particles and cell design but 100 cells each, so the sample stays
    within the time budget.  steps/s is then scaled to particles/s through the
    benchmark's measured mean steps per 2,500-cell particle."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    threads = min(os.cpu_count() or 1, 16)
    small = dict(design, cells=100)
    od = O.OracleDesign(small)
    done_steps, done_p, t_used = 0, 0, 0.0
    it = 0
    while t_used < budget_s and it < 5000:
        m = models[it % len(models)]
        first = (it // len(models)) * 16
        th = O.sample_prior(od, m, 0, first, 16)
        t = time.perf_counter()
        r = O.simulate(od, m, 0, first, th, n_threads=threads, cells=False)
        t_used += time.perf_counter() - t
        done_steps += int(r["steps"].sum())
        done_p += 16
        it += 1
    sps = done_steps / t_used
    return {"value": round(sps / gpu_steps_per_particle, 4), "unit": "particles/s",
            "cores": threads, "kind": "port",
            "sample": f"{done_p} benchmark particles (16 per model round-robin, first of each "
                      f"model's stream) with 100 cells each: {done_steps} SSA steps in "
                      f"{t_used:.1f}s on {threads} threads = {sps:.3e} steps/s; scaled by the "
                      f"benchmark's mean {gpu_steps_per_particle:.4g} steps per 2500-cell particle",
            "steps_per_s": sps}


if __name__ == "__main__":
    main()
