"""Benchmark: ABC particles/sec for the plain path (SSA + statistics + errors
+ acceptance).

Workload: the full ABC sweep (5e6 particles per model of the five models,
2,000 synthetic genes, per-gene top-1,000 acceptance), sampled: each step
takes P particles per model of that sweep.  Cells: 2,500 per particle,
theta ~ U[0,1), 21 h cycle, 5 age bins, 4 pulses + 5 chases, Beta(2,8)
capture, config priors, Philox seed 0.  Every particle's 90 superset
statistics (mean and Fano factor per age bin and condition) are scored
against genes whose data statistics come from ground-truth particles of the
same priors with SE = |stat| * U[0.05, 0.3].  One step = one batch through
prior -> rates -> burn-in -> SSA -> statistics -> split-BF16 MFMA error
screen -> FP64 rescoring of the pairs within 2 % (plus the proven bound) of
each gene's threshold -> on-device per-gene top-1,000 selection, all in HBM;
the timed region ends with the sweep's accepted sets on the host.

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


def host_info(threads):
    """The host the CPU baseline runs on: core counts, CPU model and clock, and
    the oracle's compiler and flags."""
    import subprocess
    info = {"threads_used": threads, "nproc": os.cpu_count(),
            "affinity_cpus": len(os.sched_getaffinity(0))}
    try:
        info["cgroup_cpu_max"] = open("/sys/fs/cgroup/cpu.max").read().strip()
    except OSError:
        info["cgroup_cpu_max"] = None
    model, mhz = None, None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name") and model is None:
                model = line.split(":", 1)[1].strip()
            if line.startswith("cpu MHz") and mhz is None:
                mhz = float(line.split(":", 1)[1])
    except OSError:
        pass
    info["cpu_model"], info["cpu_mhz"] = model, mhz
    try:
        info["compiler"] = subprocess.run(["gcc", "--version"], capture_output=True,
                                          text=True).stdout.splitlines()[0]
    except (OSError, IndexError):
        info["compiler"] = None
    flags = None
    for line in open(os.path.join(ROOT, "oracle", "Makefile")):
        if line.startswith("CFLAGS"):
            flags = line.split("=", 1)[1].strip().rstrip("\\").strip()
    info["flags"] = flags
    return info


def cpu_threads():
    """Every CPU this process may run on: the affinity set, bounded by a
    cgroup CPU quota when one is set."""
    n = len(os.sched_getaffinity(0))
    try:
        quota, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if quota != "max":
            n = min(n, max(1, int(int(quota) // int(period))))
    except (OSError, ValueError):
        pass
    return n


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--particles", type=int, default=16384, help="particles per model per step")
    ap.add_argument("--genes", type=int, default=2000)
    ap.add_argument("--topk", type=int, default=1000, help="accepted particles per gene")
    ap.add_argument("--margin", type=float, default=0.02,
                    help="FP64 rescoring band around each gene's threshold (relative)")
    ap.add_argument("--sweep-per-model", type=int, default=5_000_000,
                    help="particles per model of the sweep the steps sample")
    ap.add_argument("--cells", type=int, default=2500)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the occupancy-default SSA grid")
    ap.add_argument("--ssa-claim", type=int, default=0, choices=[0, 1, 2, 4, 8],
                    help="SSA work items per queue claim (0 = the autotuner's or the default)")
    ap.add_argument("--no-graph", action="store_true", help="skip the HIP-graph replay of the steps")
    ap.add_argument("--sur-train", type=int, default=200000, help="training particles per model")
    ap.add_argument("--sur-eval", type=int, default=40000, help="held-out particles per model")
    ap.add_argument("--sur-steps", type=int, default=15000, help="Adam steps (batch 8192)")
    ap.add_argument("--sur-cal", type=int, default=40000,
                    help="calibration particles per model that set the calibrated pass rules")
    ap.add_argument("--sur-cal-recall", type=float, default=0.992,
                    help="recall of the calibration set's accepted pairs each rule keeps")
    ap.add_argument("--sur-direct-steps", type=int, default=12000,
                    help="Adam steps of the network trained on accepting-gene counts (0 = skip)")
    ap.add_argument("--ceiling-reps", type=int, default=8,
                    help="independent exact replicates per held-out particle for the bound on "
                         "what any parameter-only screen can gain (0 = skip)")
    ap.add_argument("--ceiling-particles", type=int, default=10000,
                    help="held-out particles per model (the first ones) the replicates cover "
                         "(0 = all)")
    ap.add_argument("--no-moments", action="store_true",
                    help="skip the leg of the reference's own algorithm (moment model)")
    ap.add_argument("--mom-particles", type=int, default=32768,
                    help="particles per model of the moment-model leg")
    ap.add_argument("--mom-method", type=int, default=1, choices=[0, 1],
                    help="moment-model integration: 1 = gene block in closed form, "
                         "0 = Dormand-Prince on all nine moments")
    ap.add_argument("--no-reference-shape", action="store_true",
                    help="skip the leg at the reference's own design and 3,419 genes")
    ap.add_argument("--ref-particles", type=int, default=2048,
                    help="particles per model per step of the reference-shape leg")
    ap.add_argument("--ref-steps", type=int, default=2)
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

    data_f = np.asfortranarray(data)
    se_f = np.asfortranarray(se)
    S_PER = a.sweep_per_model

    def begin():
        # top-k acceptance session over the genes: per gene the k particles of
        # smallest FP64 error (ties by index) among every particle added
        _lib.check(L.abc_accept_begin(eng.ctx, KS, G, data_f.ctypes.data_as(C.c_void_p),
                                      se_f.ctypes.data_as(C.c_void_p),
                                      parts_c.ctypes.data_as(C.c_void_p), 2, 10.0, 1, 0.0,
                                      a.topk, a.margin), eng.ctx, "accept_begin")

    def end():
        """Close the session: per-gene counts, 1-based indices, errors and the
        number of FP64 rescorings."""
        cap = G * a.topk + 16
        counts = np.zeros(G, np.int64)
        idx = np.zeros(cap, np.int64)
        eo = np.zeros(cap)
        nc = C.c_int64()
        _lib.check(L.abc_accept_end(eng.ctx, counts.ctypes.data_as(C.c_void_p),
                                    idx.ctypes.data_as(C.c_void_p), eo.ctypes.data_as(C.c_void_p),
                                    cap, C.byref(nc)), eng.ctx, "accept_end")
        return counts, idx, eo, nc.value

    acc_ev = []

    def step(it, timed):
        first = it * P
        for m in models:
            _lib.check(L.abc_sample_prior_dev(eng.ctx, m, 0, first, P, ptr(bufs[m]["theta"])),
                       eng.ctx, "prior")
        _lib.check(L.abc_simulate_multi_dev(eng.ctx, nm, m_models, 0, (C.c_int64 * nm)(*([first] * nm)),
                                            m_n, m_theta, m_stats, None, m_sup, m_steps),
                   eng.ctx, "simulate")
        if timed:
            e0 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
        for q, m in enumerate(models):
            b = bufs[m]
            # particle `first + i` of model m is sweep particle q * S_PER + first + i
            _lib.check(L.abc_accept_add_dev(eng.ctx, P, ptr(b["sup"]), q * S_PER + first),
                       eng.ctx, "accept_add")
            if timed:
                steps_acc.add_(b["steps"].sum())  # same stream: no host sync
        if timed:
            e1 = torch.cuda.Event(enable_timing=True)
            e1.record(stream)
            acc_ev.append((e0, e1))

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
        print(f"[bench] autotune: {blocks} workgroups, {eng.ssa_claim()} items per claim",
              file=sys.stderr, flush=True)
    if a.ssa_claim:
        eng.set_ssa_claim(a.ssa_claim)
    begin()
    for w in range(a.warmup):
        step(10_000 + w, False)
    end()
    torch.cuda.synchronize()
    ssa_time(1)
    begin()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    for it in range(a.steps):
        step(it, True)
        print(f"[bench] step {it} enqueued", file=sys.stderr, flush=True)
    counts, idx, eo, n_resc = end()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    accept_ms = sum(e0.elapsed_time(e1) for e0, e1 in acc_ev)
    gpu_ms = ev0.elapsed_time(ev1)
    n_particles = a.steps * P * len(models)
    value = n_particles / (gpu_ms / 1e3)
    burn_ms_total, ssa_ms_total, burn_steps, launches = ssa_time(1)
    total_steps = int(steps_acc.item())
    window_steps = total_steps - burn_steps

    graph = None
    if not a.no_graph:
        # the first steps again, each captured into one HIP graph (its own
        # acceptance session; the graph is updated in place for the next step's
        # particles) and replayed; only the replays are timed
        gsteps = min(a.steps, 2)
        begin()
        g, gms = None, 0.0
        for it in range(gsteps):
            g = eng.capture(lambda: step(it, False), graph=g)
            ev2, ev3 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            ev2.record(stream)
            eng.replay(g)
            ev3.record(stream)
            torch.cuda.synchronize()
            gms += ev2.elapsed_time(ev3)
        end()
        eng.drop_graph(g)
        ssa_time(1)
        graph = {"value": round(gsteps * P * len(models) / (gms / 1e3), 3),
                 "ms_per_step": round(gms / gsteps, 3), "steps": gsteps,
                 "timed": "graph launches only (capture and update outside the timing)"}
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
        "config": {"workload": "full ABC sweep (5 models x %d particles, %d genes, top-%d per gene), "
                               "sampled: %d of its %d particles (%d per model per step)"
                               % (S_PER, G, a.topk, n_particles, S_PER * len(models), P),
                   "models": models, "particles_per_model_per_step": P, "cells_per_particle": a.cells,
                   "genes": G, "stats": KS, "cycle_h": 21.0, "capture": "Beta(2,8)",
                   "acceptance": {"rule": "top-k per gene", "k": a.topk,
                                  "rescoring_band": a.margin,
                                  "fp64_rescorings": n_resc,
                                  "pairs": n_particles * G,
                                  "accepted_pairs": int(counts.sum()),
                                  "ms_per_step": round(accept_ms / a.steps, 3)},
                   "mean_steps_per_particle": total_steps / n_particles},
        "roofline": roof,
        "wall_s": round(wall, 3),
        "ssa_launch": {"workgroups": eng.ssa_blocks(), "wg_lanes": 256,
                       "items_per_claim": eng.ssa_claim(), "autotune": tuned},
        "graph_replay": graph,
    }

    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(design, models, P, total_steps / n_particles, a.cpu_seconds)
    if not a.no_surrogate:
        out["surrogate"] = surrogate_leg(eng, stream, a)
    if not a.no_moments:
        out["reference_moment_path"] = moment_leg(stream, a)
    if not a.no_reference_shape:
        out["reference_shape"] = reference_shape_leg(stream, a)
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
                accept_rate=4e-5, z=[3.0, 2.0], train_steps=a.sur_steps, batch=8192,
                ceiling_reps=a.ceiling_reps, n_cal=a.sur_cal, cal_recall=a.sur_cal_recall,
                direct_steps=a.sur_direct_steps, ceiling_particles=a.ceiling_particles,
                log=lambda *x: print(*x, file=sys.stderr, flush=True))
    sur.close()
    keep = ("plain_accepted_per_s", "screened_accepted_per_s", "speedup_accepted_per_s",
            "recall", "ssa_fraction_avoided", "ssa_work_fraction_avoided", "plain_pairs",
            "eval_particles", "train_particles", "heldout_nll", "heldout_nll_of_mean_predictor",
            "heldout_coverage_2sigma", "heldout_z_rms", "heldout_rmse_std_units_median",
            "train_loss_last", "train_s", "direct_train_s", "train_simulate_s", "plain_time_s",
            "plain_simulate_time_s", "heldout_coverage_2sigma_per_stat",
            "heldout_rmse_std_units", "heldout_rmse_std_units_max", "genes", "z", "eps", "accept_rate_target", "by_z",
            "ceiling", "by_rule", "rule", "threshold")
    return {k: res[k] for k in keep if k in res}


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


def reference_shape_leg(stream, a):
    """The drop-in at the reference's own shape (wrapper.jl:59-78): T = 20 h,
    ages at the bin centres, 11 conditions, capture from the shipped
    efficiencies, the 53 abc_sim statistics scored against all 3,419 genes of
    data/summary_stats and accepted at the fixed epsilon 4.8
    (scripts/accepted_particles.jl:10); 5,445 cells per particle.  Each step
    simulates P particles of each of the five models (reference prior) in one
    launch and adds them to one acceptance session (indices offset per model,
    so each gene's list splits into the per-model files); the timed region
    ends with the accepted sets on the host."""
    import torch
    from abc_inference_transcription_amd import _lib
    from abc_inference_transcription_amd.design import reference_shape_design
    from abc_inference_transcription_amd.engine import Engine
    from abc_inference_transcription_amd.host import data_matrices
    golden = os.path.join(ROOT, "tests", "golden")
    betas = np.load(os.path.join(golden, "capture.npz"))["betas"]
    data, se = data_matrices(dict(np.load(os.path.join(golden, "summary_stats_all.npz"))))
    data, se = np.asfortranarray(data), np.asfortranarray(se)
    G, K = data.shape
    eng = Engine(reference_shape_design(betas), device=0)
    eng.bind_stream(stream)
    L = eng.L
    dev = torch.device("cuda:0")
    models, P, per = [1, 2, 3, 4, 5], a.ref_particles, 250_000
    nm = len(models)
    th = [torch.empty(eng.n_params(m) * P, dtype=torch.float64, device=dev) for m in models]
    st = [torch.empty(K * P, dtype=torch.float64, device=dev) for _ in models]
    VA = C.c_void_p * nm
    m_models = (C.c_int32 * nm)(*models)
    m_n = (C.c_int64 * nm)(*([P] * nm))
    m_th = VA(*[t.data_ptr() for t in th])
    m_st = VA(*[t.data_ptr() for t in st])
    parts = np.ascontiguousarray(eng.parts, np.int32)
    vp = lambda x: x.ctypes.data_as(C.c_void_p)

    def step(it):
        first = it * P
        for q, m in enumerate(models):
            _lib.check(L.abc_sample_prior_dev(eng.ctx, m, 0, first, P,
                                              C.c_void_p(th[q].data_ptr())), eng.ctx, "prior")
        _lib.check(L.abc_simulate_multi_dev(eng.ctx, nm, m_models, 0,
                                            (C.c_int64 * nm)(*([first] * nm)), m_n, m_th, m_st,
                                            None, None, None), eng.ctx, "simulate")
        for q in range(nm):
            _lib.check(L.abc_accept_add_dev(eng.ctx, P, C.c_void_p(st[q].data_ptr()),
                                            q * per + first), eng.ctx, "accept_add")

    def session(steps, base):
        _lib.check(L.abc_accept_begin(eng.ctx, K, G, vp(data), vp(se), vp(parts), len(parts),
                                      10.0, 0, 4.8, 0, 0.02), eng.ctx, "accept_begin")
        for it in range(steps):
            step(base + it)
        cap = 1 << 22
        while True:  # a short capacity leaves the session open with the counts filled in
            counts = np.zeros(G, np.int64)
            idx = np.zeros(cap, np.int64)
            nc = C.c_int64()
            rc = L.abc_accept_end(eng.ctx, vp(counts), vp(idx), None, cap, C.byref(nc))
            if rc != -5:
                break
            cap = int(counts.sum())
        _lib.check(rc, eng.ctx, "accept_end")
        return counts, idx[:int(counts.sum())], nc.value

    session(1, 1000)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    counts, idx, nres = session(a.ref_steps, 0)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    n = a.ref_steps * P * nm
    per_model = [int(((idx - 1) // per == q).sum()) for q in range(nm)]
    eng.close()
    res = {"value": round(n / wall, 3), "unit": "particles/s", "particles": n,
           "cells_per_particle": 5445, "genes": G, "stats": K, "eps": 4.8,
           "accepted_pairs": int(counts.sum()), "accepted_pairs_per_model": per_model,
           "genes_with_acceptance": int((counts > 0).sum()), "fp64_rescorings": nres,
           "ms_per_step": round(wall / a.ref_steps * 1e3, 3),
           "timing": "host clock around the session (simulation, screening, FP64 rescoring, "
                     "per-batch merge, sorted lists on the host)"}
    print(f"[bench] reference shape: {res['value']:.1f} particles/s, "
          f"{res['accepted_pairs']} accepted pairs", file=sys.stderr, flush=True)
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
        threads = cpu_threads()
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
    threads = cpu_threads()
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
            "kind": "port", "host": host_info(threads), "sample": f"{done} particles (16 per model round-robin) in "
                                      f"{t_used:.1f}s on {threads} threads"}


def cpu_baseline(design, models, P, gpu_steps_per_particle, budget_s):
    """The oracle's scalar replay of the same SSA (the identical Philox steps)
    on the host cores, over a bounded sample of the benchmark's own particles
    (the first ones of each model's stream, the benchmark's design and cells,
    cells spread over the threads), for about budget_s seconds.  Particle
    costs are heavy-tailed, so the sample's steps/s is scaled to particles/s
    through the benchmark's measured mean steps per particle; the sample's
    own particles/s is reported beside it."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    threads = cpu_threads()
    od = O.OracleDesign(design)
    cells = int(design["cells"])
    batch = 8
    done_steps, done_p, t_used = 0, 0, 0.0
    it = 0
    while t_used < budget_s and it < 5000:
        m = models[it % len(models)]
        first = (it // len(models)) * batch
        th = O.sample_prior(od, m, 0, first, batch)
        t = time.perf_counter()
        r = O.simulate(od, m, 0, first, th, n_threads=threads, cells=False)
        t_used += time.perf_counter() - t
        done_steps += int(r["steps"].sum())
        done_p += batch
        it += 1
    sps = done_steps / t_used
    return {"value": round(sps / gpu_steps_per_particle, 4), "unit": "particles/s",
            "cores": threads, "kind": "port", "host": host_info(threads),
            "sample": f"{done_p} benchmark particles ({batch} per model round-robin, first of each "
                      f"model's stream) with {cells} cells each: {done_steps} SSA steps in "
                      f"{t_used:.1f}s on {threads} threads = {sps:.3e} steps/s "
                      f"({done_p / t_used:.2f} particles/s on this sample); scaled by the "
                      f"benchmark's mean {gpu_steps_per_particle:.4g} steps per particle",
            "steps_per_s": sps, "sample_particles_per_s": done_p / t_used}


if __name__ == "__main__":
    main()
