"""Surrogate-screened ABC against the exhaustive plain path.

1. Synthetic genes: data statistics of ground-truth particles drawn from the
   same priors, SE = |stat| * U[0.05, 0.3]; genes whose condition-averaged
   mean never reaches 1 are dropped (the expression filter of the gene
   selection step).
2. Training set: exact simulations of prior particles (all five models); the
   surrogate learns mean and log-variance of each standardised asinh(stat)
   under a Gaussian NLL.
3. Acceptance rule: fixed epsilon (k == 0), calibrated so that a target
   fraction ``accept_rate`` of (training particle, gene) pairs falls below it,
   or per-gene top-k (k > 0).
4. Plain path on held-out particles: exact simulation of every particle and
   FP64 acceptance of every pair -> accepted pairs A_plain.
5. Screened path on the same particles, once per z: the surrogate predicts
   each statistic's distribution, the screen kernel bounds each gene's error
   from below (mean - z * sd of the error under the predicted moments), and
   only particles that can pass some gene are simulated exactly and accepted
   in FP64 -> A_screened.
6. Calibrated rules (n_cal > 0): scores of particles (expected accepting
   genes and largest margin from the statistics surrogate; with
   direct_steps > 0, a second network's expected accepting genes and expected
   genes per unit of exact cost) with pass thresholds set on a separate,
   exactly simulated calibration set at the target recall, then applied to
   the held-out particles like the z rule.
Reports recall |A_plain & A_screened| / |A_plain|, the fraction of exact
simulation work avoided and accepted pairs per second of both paths, per z
(``by_z``), per calibrated rule (``by_rule``) and for the fastest setting
that keeps recall >= 0.99.
All compute runs in libabc.so; torch only holds device buffers.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import _lib

MODELS = (1, 2, 3, 4, 5)


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def simulate_block(eng, model, n, first, seed=0, with_steps=False):
    """Prior draws and exact simulation of particles first..first+n-1 on the
    device: (theta, superset n x K column-major, steps)."""
    import torch
    dev = torch.device("cuda:0")
    L = eng.L
    npar = eng.n_params(model)
    KS = 2 * eng.nb * eng.J
    th = torch.empty(npar * n, dtype=torch.float64, device=dev)
    st = torch.empty((4 * eng.nb + 3 * eng.J) * n, dtype=torch.float64, device=dev)
    sup = torch.empty(KS * n, dtype=torch.float64, device=dev)
    steps = torch.zeros(n, dtype=torch.int64, device=dev)
    _lib.check(L.abc_sample_prior_dev(eng.ctx, model, seed, first, n, _ptr(th)), eng.ctx, "prior")
    _lib.check(L.abc_simulate_dev(eng.ctx, model, seed, first, n, _ptr(th), _ptr(st), None,
                                  _ptr(sup), _ptr(steps)), eng.ctx, "simulate")
    return th, sup, steps


def simulate_models(eng, models, n, firsts, seed=0):
    """Prior draws of n particles per model (particles firsts[m] ... of model
    m) and their exact simulation in one multi-model call: m -> (theta,
    superset n x K column-major, steps), all on the device."""
    import torch
    dev = torch.device("cuda:0")
    L = eng.L
    KS = 2 * eng.nb * eng.J
    out = {}
    for m in models:
        th = torch.empty(eng.n_params(m) * n, dtype=torch.float64, device=dev)
        _lib.check(L.abc_sample_prior_dev(eng.ctx, m, seed, firsts[m], n, _ptr(th)), eng.ctx,
                   "prior")
        out[m] = (th, torch.empty((4 * eng.nb + 3 * eng.J) * n, dtype=torch.float64, device=dev),
                  torch.empty(KS * n, dtype=torch.float64, device=dev),
                  torch.zeros(n, dtype=torch.int64, device=dev))
    k = len(models)
    VA = C.c_void_p * k
    _lib.check(L.abc_simulate_multi_dev(
        eng.ctx, k, (C.c_int32 * k)(*models), seed, (C.c_int64 * k)(*[firsts[m] for m in models]),
        (C.c_int64 * k)(*([n] * k)), VA(*[out[m][0].data_ptr() for m in models]),
        VA(*[out[m][1].data_ptr() for m in models]), None,
        VA(*[out[m][2].data_ptr() for m in models]), VA(*[out[m][3].data_ptr() for m in models])),
        eng.ctx, "simulate_multi")
    return {m: (th, sup, stp) for m, (th, _, sup, stp) in out.items()}


def make_genes(eng, G, seed=0, first=3_000_000_000):
    """Data statistics of G ground-truth prior particles that pass the
    reference's expression filter (scripts/gene_selection.jl:13-25: on average
    at least one count per cell in some age bin); rows with non-finite
    statistics (an age/condition group with fewer than two cells) are
    skipped."""
    nb, J = eng.nb, eng.J
    rows, got, block = [], 0, 4 * G
    while got < G:
        per = (block + len(MODELS) - 1) // len(MODELS)
        for m in MODELS:
            _, sup, _ = simulate_block(eng, m, per, first + m * per, seed)
            KS = sup.numel() // per
            r = sup.view(KS, per).t().cpu().numpy()
            means = r[:, :nb * J].reshape(-1, J, nb)          # mean_t(b, j) at b + nb*j
            per_bin = means.mean(axis=1)                      # average over conditions
            ok = np.all(np.isfinite(r), axis=1) & (per_bin.max(axis=1) >= 1.0)
            rows.append(r[ok])
            got += int(ok.sum())
        first += len(MODELS) * per + per
        if first > 4_200_000_000:
            raise RuntimeError("not enough expressed ground-truth particles for the genes")
    data = np.vstack(rows)
    order = np.random.default_rng(seed).permutation(len(data))
    data = data[order][:G]
    if len(data) < G:
        raise RuntimeError("not enough finite ground-truth particles for the genes")
    rng = np.random.default_rng(seed)
    se = np.abs(data) * rng.uniform(0.05, 0.3, data.shape)
    return np.asfortranarray(data), np.asfortranarray(se)


def run(eng, sur, G=2000, n_train=6000, n_eval=2000, k=0, eps=4.8, accept_rate=4e-5,
        z=4.0, margin=0.0,
        train_steps=1500, batch=8192, lr=1e-3, seed=0, ceiling_reps=0, n_cal=0,
        cal_recall=0.995, direct_steps=0, direct_loss="poisson", fixed_share=0.5, direct_nets=2,
        ceiling_particles=0, direct_aux=(1.5, 3.0), log=print):
    """Returns a dict of metrics (see module docstring)."""
    import torch
    dev = torch.device("cuda:0")
    # one stream for libabc, the surrogate and torch's own buffer work, so every
    # kernel sees the outputs of the ones enqueued before it
    stream = torch.cuda.current_stream()
    if stream.cuda_stream == 0:
        stream = torch.cuda.Stream(device=dev)
        torch.cuda.set_stream(stream)
    eng.bind_stream(stream)
    sur.bind_stream(stream)
    L = eng.L
    nb, J = eng.nb, eng.J
    KS = 2 * nb * J
    parts = np.array([nb * J, nb * J], np.int32)
    mode = 1 if k > 0 else 0   # top-k per gene, or the reference's fixed epsilon
    out = {}
    t0 = time.perf_counter()
    data, se = make_genes(eng, G, seed)
    log(f"[screened] genes ready ({time.perf_counter() - t0:.1f}s)")

    # ---- training set (streams 1e9 + ...), disjoint from the evaluation set
    xs, ts_raw, ts_steps = [], [], []
    t0 = time.perf_counter()
    for m in MODELS:
        th, sup, stp = simulate_block(eng, m, n_train, 1_000_000_000 + m * n_train, seed)
        xs.append(sur.features(m, th))
        ts_raw.append(sup)
        ts_steps.append(stp)
    torch.cuda.synchronize()
    out["train_simulate_s"] = time.perf_counter() - t0
    stats_np = np.vstack([s.view(KS, n_train).t().cpu().numpy() for s in ts_raw])
    sur.fit_normalizer(stats_np)
    x_train = torch.cat(xs)
    t_train = torch.cat([sur.targets(s, n_train) for s in ts_raw])
    t0 = time.perf_counter()
    losses = sur.train(x_train, t_train, steps=train_steps, batch=batch, lr=lr, seed=seed)
    out["train_s"] = time.perf_counter() - t0
    out["train_loss_first"] = float(losses[:10].mean())
    out["train_loss_last"] = float(losses[-50:].mean())
    log(f"[screened] trained {train_steps} steps: loss {out['train_loss_first']:.3f} -> "
        f"{out['train_loss_last']:.3f} ({out['train_s']:.1f}s)")

    # ---- acceptance threshold at the target selectivity: the FP64 errors of
    # the (already simulated) training particles against the genes give the
    # epsilon at which a fraction `accept_rate` of (particle, gene) pairs pass
    # (a sweep that keeps the best 1,000 of 2.5e7 particles per gene accepts 4e-5)
    if accept_rate is not None and mode == 0:
        errs = []
        data_d = torch.tensor(data.ravel(order="F"), dtype=torch.float64, device=dev)
        se_d = torch.tensor(se.ravel(order="F"), dtype=torch.float64, device=dev)
        ne = min(n_train, 40_000)   # the first particles of each model: 4e8 pairs at most
        for s in ts_raw:
            e = torch.empty(ne * G, dtype=torch.float64, device=dev)
            blk = s.view(KS, n_train)[:, :ne].contiguous()
            _lib.check(L.abc_errors_dev(eng.ctx, ne, KS, _ptr(blk), G, _ptr(data_d), _ptr(se_d),
                                        parts.ctypes.data_as(C.c_void_p), 2, 10.0, _ptr(e)),
                       eng.ctx, "errors")
            errs.append(e)
        e = torch.cat(errs)
        e = e[torch.isfinite(e)]
        kth = max(1, int(round(accept_rate * e.numel())))
        eps = float(torch.kthvalue(e, kth).values.item())
        out["accept_rate_target"] = accept_rate
        log(f"[screened] epsilon {eps:.4g} accepts {accept_rate:g} of training pairs")
    out["eps"] = eps if mode == 0 else None
    direct = None
    if direct_steps and mode == 0:
        direct = _train_direct(eng, sur, L, data, se, parts, KS, G, eps, xs, ts_raw, ts_steps,
                               n_train, direct_steps, batch, lr, seed, dev, log, direct_loss,
                               fixed_share, direct_nets, direct_aux)
        out["direct_train_s"] = direct.train_s

    # ---- exhaustive plain path on the evaluation set (timed)
    def eval_first(m):
        return 2_000_000_000 + m * n_eval

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _lib.check(L.abc_accept_begin(eng.ctx, KS, G, data.ctypes.data_as(C.c_void_p),
                                  se.ctypes.data_as(C.c_void_p),
                                  parts.ctypes.data_as(C.c_void_p), 2, 10.0, mode, eps, k, 0.02),
               eng.ctx, "accept_begin")
    eval_sets, eval_steps = {}, {}
    sims = simulate_models(eng, MODELS, n_eval, {m: eval_first(m) for m in MODELS}, seed)
    torch.cuda.synchronize()
    t_plain_sim = time.perf_counter() - t0
    for m in MODELS:
        th, sup, steps = sims[m]
        eval_sets[m] = (th, sup)
        eval_steps[m] = steps
        _lib.check(L.abc_accept_add_dev(eng.ctx, n_eval, _ptr(sup), eval_first(m)), eng.ctx, "add")
    A_plain = _end(eng, G, G * max(k, 16))
    t_plain = time.perf_counter() - t0
    n_all = n_eval * len(MODELS)
    thr = (np.array([max(e) if len(e) else 10.0 for (_, e) in A_plain]) if mode == 1
           else np.full(G, eps))
    plain_pairs = {(g, int(i)) for g, (idx, _) in enumerate(A_plain) for i in idx}
    out.update(plain_time_s=t_plain, plain_simulate_time_s=t_plain_sim,
               plain_pairs=len(plain_pairs), eval_particles=n_all,
               plain_accepted_per_s=len(plain_pairs) / t_plain)
    log(f"[screened] plain: {len(plain_pairs)} accepted pairs in {t_plain:.2f}s (exact "
        f"simulation {t_plain_sim:.3f}s)")

    # ---- held-out accuracy of the surrogate against the simulator (per statistic,
    # in standardised asinh units; the targets carry the sampling noise of the
    # ~55 cells behind each (bin, condition) statistic, which the predicted
    # variance is meant to capture)
    err_rms, cover, nll, zsq = [], [], [], []
    for m in MODELS:
        th, sup = eval_sets[m]
        x = sur.features(m, th)
        y = sur.predict(x)
        t = sur.targets(sup, n_eval)
        lv = y[:, KS:].clamp(-12, 12)
        mu, sd = y[:, :KS], torch.exp(0.5 * lv)
        err_rms.append(((mu - t) ** 2).mean(0))
        cover.append(((mu - t).abs() <= 2 * sd).float().mean(0))
        nll.append((0.5 * (lv + (t - mu) ** 2 * torch.exp(-lv))).mean(0))
        zsq.append((((t - mu) / sd) ** 2).mean(0))
    err_rms = torch.sqrt(torch.stack(err_rms).mean(0)).cpu().numpy()
    cover = torch.stack(cover).mean(0).cpu().numpy()
    out["heldout_rmse_std_units"] = [round(float(v), 5) for v in err_rms]
    out["heldout_coverage_2sigma_per_stat"] = [round(float(v), 4) for v in cover]
    out["heldout_rmse_std_units_median"] = float(np.median(err_rms))
    out["heldout_rmse_std_units_max"] = float(err_rms.max())
    out["heldout_coverage_2sigma"] = float(cover.mean())
    out["heldout_nll"] = float(torch.stack(nll).mean().item())
    out["heldout_z_rms"] = float(np.sqrt(torch.stack(zsq).mean().item()))
    out["heldout_nll_of_mean_predictor"] = 0.5  # N(0,1) on standardised targets

    # ---- screened path (timed) for each z: surrogate screen, then exact work
    # on the particles that pass only
    zs = list(z) if isinstance(z, (list, tuple)) else [z]
    by_z = []
    for zz in zs:
        by_z.append(_screened(eng, sur, L, data, se, parts, KS, G, mode, eps, k, thr, zz, margin,
                              eval_sets, eval_steps, eval_first, n_eval, n_all, plain_pairs,
                              t_plain, seed, dev, log))
    out["by_z"] = by_z
    if n_cal and mode == 0:
        out["by_rule"] = _calibrated_rules(eng, sur, L, data, se, parts, KS, G, eps, n_cal,
                                           cal_recall, eval_sets, eval_steps, eval_first,
                                           n_eval, n_all, plain_pairs, t_plain, seed, dev, log,
                                           direct=direct)
        by_z = by_z + out["by_rule"]
    if ceiling_reps and mode == 0:
        # the first ceiling_particles of each model's held-out set (0: all)
        nc = min(n_eval, ceiling_particles) if ceiling_particles > 0 else n_eval
        th_c = {m: eval_sets[m][0].view(-1, n_eval)[:, :nc].contiguous().view(-1) for m in MODELS}
        out["ceiling"] = acceptance_ceiling(
            eng, data, se, parts, eps, th_c, {m: eval_first(m) for m in MODELS},
            reps=ceiling_reps, log=log)
        log(f"[screened] ceiling: {out['ceiling']}")
    # the headline: calibrated rules when there are any (their thresholds never
    # saw the evaluation particles), else the best z of the grid
    pool = out.get("by_rule") or by_z
    best = [r for r in pool if r["recall"] >= 0.99]
    pick = max(best, key=lambda r: r["speedup_accepted_per_s"]) if best else pool[0]
    out.update(pick)
    out.update(genes=G, train_particles=n_train * len(MODELS), k=k)
    if direct is not None:
        direct.close()
    return out


def _accepting_genes(eng, L, sup, n, data_d, se_d, parts, KS, G, eps, dev, chunk=50_000):
    """Per particle (statistics sup: n x KS on the device), the number of genes
    whose FP64 error is <= eps (torch int32, device)."""
    import torch
    out = torch.empty(n, dtype=torch.int32, device=dev)
    err = torch.empty(min(n, chunk) * G, dtype=torch.float64, device=dev)
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        # rows a..b-1 of the column-major n x KS matrix: a block with leading dimension n
        blk = sup.view(KS, n)[:, a:b].contiguous()
        _lib.check(L.abc_errors_dev(eng.ctx, b - a, KS, _ptr(blk), G, _ptr(data_d), _ptr(se_d),
                                    parts.ctypes.data_as(C.c_void_p), len(parts), 10.0,
                                    _ptr(err)), eng.ctx, "errors")
        out[a:b] = (err[:(b - a) * G].view(G, b - a) <= eps).sum(0).to(torch.int32)
    return out


class _Direct:
    """A network trained on the number of genes that accept each training
    particle (one noisy draw per particle) and on the log of its exact
    simulation steps: its outputs estimate E[accepting genes | parameters]
    (Poisson NLL on a log-rate output, or the Gaussian NLL's mean) and
    E[log steps | parameters] (Gaussian NLL).  Passing particles in decreasing
    expected count keeps the most accepted pairs per particle simulated; in
    decreasing count per unit of expected cost, the most per unit of exact
    work.  A particle's cost is its steps plus c0, the fixed cost of a
    particle in step units (its cells' set-up, burn-in records, queue refills
    and statistics, which do not scale with its steps).  Several networks
    (trained from different initial weights and batch orders) average their
    outputs.  Trained for one gene set."""

    def __init__(self, nets, scale, lmu, lsd, train_s, poisson, c0, n_aux=0):
        self.nets, self.scale, self.lmu, self.lsd = nets, scale, lmu, lsd
        self.train_s, self.poisson, self.c0 = train_s, poisson, c0
        self.n_aux = n_aux

    def close(self):
        for n in self.nets:
            n.close()

    def _y(self, x):
        """Averaged (count output, standardised log steps) per particle."""
        cols = [0, 1 + self.n_aux]
        y = self.nets[0].predict(x)[:, cols]
        for n in self.nets[1:]:
            y = y + n.predict(x)[:, cols]
        return (y / len(self.nets)).contiguous()

    def _count(self, y):
        import torch
        return torch.exp(y[:, 0]) if self.poisson else y[:, 0] * self.scale

    def score(self, x):
        return self._count(self._y(x))

    def per_work(self, x, other=None, alpha=0.75):
        """Expected accepting genes per unit of expected cost; with `other`
        (another estimate of the expected genes, the statistics surrogate's),
        the genes are the weighted geometric mean count^alpha other^(1-alpha)."""
        import torch
        y = self._y(x)
        cnt = self._count(y)
        if other is not None:
            cnt = cnt ** alpha * other.clamp(min=1e-12) ** (1.0 - alpha)
        return cnt / (torch.exp(y[:, 1] * self.lsd + self.lmu) + self.c0)


def _train_direct(eng, sur, L, data, se, parts, KS, G, eps, xs, ts_raw, ts_steps, n_train, steps,
                  batch, lr, seed, dev, log, loss="poisson", fixed_share=0.5, n_nets=2,
                  aux=(1.5, 3.0)):
    """Count networks: outputs (accepting genes at eps, at eps * a for each a
    in aux, standardised log steps).  The auxiliary counts at looser
    thresholds are denser targets (about 4x and 7x as many training
    particles with a count at 1.5 and 3 eps) that shape the shared hidden layers; only
    the count at eps and the steps are read (Poisson counts only)."""
    import torch
    from .surrogate import Surrogate
    data_d = torch.tensor(np.asfortranarray(data).ravel(order="F"), dtype=torch.float64,
                          device=dev)
    se_d = torch.tensor(np.asfortranarray(se).ravel(order="F"), dtype=torch.float64, device=dev)
    poisson = loss == "poisson"
    aux = tuple(aux) if poisson else ()
    cnts = [torch.cat([_accepting_genes(eng, L, s, n_train, data_d, se_d, parts, KS, G, eps * a,
                                        dev) for s in ts_raw]).to(torch.float32)
            for a in (1.0,) + aux]
    cnt = cnts[0]
    scale = 1.0 if poisson else float(max(cnt.std().item(), 1e-3))
    lw = torch.log(torch.cat(ts_steps).to(torch.float64).clamp(min=1.0)).to(torch.float32)
    lmu, lsd = float(lw.mean().item()), float(max(lw.std().item(), 1e-3))
    c0 = fixed_share * float(torch.cat(ts_steps).to(torch.float64).mean().item())
    nc = len(cnts)
    tgt = torch.stack([cnt / scale] + cnts[1:] + [(lw - lmu) / lsd], dim=1).contiguous()
    x_all = torch.cat(xs)
    nets = []
    t0 = time.perf_counter()
    for q in range(n_nets):
        net = Surrogate(nc + 1, n_bins=eng.nb, max_batch=sur.max_batch,
                        seed=seed + 17 + 1000 * q, stream=sur.stream)
        if poisson:
            net.set_count_outputs(nc)
            # start each log-rate at its mean count: small first gradients
            w = net.get_params()
            out_pad = (2 * (nc + 1) + 127) // 128 * 128
            for i in range(nc):
                w[len(w) - out_pad + i] = np.log(max(float(cnts[i].mean().item()), 1e-3))
            net.set_params(w)
        losses = net.train(x_all, tgt, steps=steps, batch=batch, lr=lr, seed=seed + 1 + 1000 * q)
        nets.append(net)
    train_s = time.perf_counter() - t0
    log(f"[screened] direct network: {int((cnt > 0).sum().item())} of {len(cnt)} training "
        f"particles accepted by some gene (auxiliary thresholds x{list(aux)}: "
        f"{[int((c > 0).sum().item()) for c in cnts[1:]]}); loss {losses[:10].mean():.3f} -> "
        f"{losses[-50:].mean():.3f} ({train_s:.1f}s)")
    return _Direct(nets, scale, lmu, lsd, train_s, poisson, c0, n_aux=len(aux))


def _calibrated_rules(eng, sur, L, data, se, parts, KS, G, eps, n_cal, cal_recall, eval_sets,
                      eval_steps, eval_first, n_eval, n_all, plain_pairs, t_plain, seed, dev,
                      log, direct=None):
    """The screen's two scores (expected accepting genes; largest standardised
    margin) on a calibration set of exactly simulated particles fix, per rule,
    the largest pass threshold whose recall of the calibration set's accepted
    pairs is at least cal_recall.  Each rule then screens the held-out
    evaluation particles with that threshold (timed like the z rule), so the
    reported recall is measured on particles the threshold never saw."""
    import torch
    data_d = torch.tensor(np.asfortranarray(data).ravel(order="F"), dtype=torch.float64,
                          device=dev)
    se_d = torch.tensor(np.asfortranarray(se).ravel(order="F"), dtype=torch.float64, device=dev)
    _lib.check(L.abc_accept_begin(eng.ctx, KS, G, data.ctypes.data_as(C.c_void_p),
                                  se.ctypes.data_as(C.c_void_p),
                                  parts.ctypes.data_as(C.c_void_p), 2, 10.0, 0, eps, 0, 0.02),
               eng.ctx, "accept_begin")
    scores, acc = [], []
    for m in MODELS:
        first = 2_500_000_000 + m * n_cal
        th, sup, _ = simulate_block(eng, m, n_cal, first, seed)
        err = torch.empty(n_cal * G, dtype=torch.float64, device=dev)
        _lib.check(L.abc_errors_dev(eng.ctx, n_cal, KS, _ptr(sup), G, _ptr(data_d), _ptr(se_d),
                                    parts.ctypes.data_as(C.c_void_p), len(parts), 10.0,
                                    _ptr(err)), eng.ctx, "errors")
        acc.append((err.view(G, n_cal) <= eps).sum(0).cpu().numpy())
        x = sur.features(m, th)
        sc = sur.scores(eng, x)
        if direct is not None:
            sc = torch.cat([sc, direct.score(x).view(-1, 1), direct.per_work(x).view(-1, 1),
                            direct.per_work(x, sc[:, 0]).view(-1, 1)], dim=1)
        scores.append(sc.cpu().numpy())
    _end(eng, G, 16)
    acc, scores = np.concatenate(acc), np.vstack(scores)
    log(f"[screened] calibration: {len(acc)} particles, {int(acc.sum())} accepted pairs")
    res = []
    rules = [(0, "expected_genes", lambda m, x: sur.scores(eng, x)[:, 0]),
             (1, "max_margin", lambda m, x: sur.scores(eng, x)[:, 1])]
    if direct is not None:
        rules.append((2, "direct_expected_genes", lambda m, x: direct.score(x)))
        rules.append((3, "direct_genes_per_step", lambda m, x: direct.per_work(x)))
        rules.append((4, "blended_genes_per_step",
                      lambda m, x: direct.per_work(x, sur.scores(eng, x)[:, 0])))
    for col, name, fn in rules:
        sc = scores[:, col]
        order = np.argsort(-sc, kind="stable")
        cum = np.cumsum(acc[order]) / max(acc.sum(), 1)
        r = int(np.searchsorted(cum, cal_recall))
        tau = float(sc[order[min(r, len(order) - 1)]])
        r = _screened(eng, sur, L, data, se, parts, KS, G, 0, eps, 0, np.full(G, eps), None, 0.0,
                      eval_sets, eval_steps, eval_first, n_eval, n_all, plain_pairs, t_plain,
                      seed, dev, log, rule=(fn, tau))
        r.update(rule=name, threshold=tau, cal_recall_target=cal_recall,
                 cal_pass_share=float((sc >= tau).mean()))
        res.append(r)
    return res


def acceptance_ceiling(eng, data, se, parts, eps, thetas, firsts, reps=16, seed0=1,
                       recalls=(0.99, 0.995), log=print):
    """What any screen that sees only a particle's parameters can gain.

    The plain path's acceptance of a pair (particle, gene) is random given the
    particle's parameters: its statistics carry the sampling noise of the
    simulated cells.  Each held-out particle (thetas[m], particles firsts[m]
    ... of model m) is simulated again `reps` times with independent random
    streams (seeds seed0, seed0 + 1, ...), which estimates p(i, g), the
    probability that the pair is accepted (FP64 error <= eps).  A particle
    carries s_i = sum_g p(i, g) expected accepted pairs; a screen that
    passes particles in decreasing s_i reaches recall R with the smallest
    share of particles, f(R).  Particles differ in cost (w_i, their mean exact
    simulation steps), and the time of the screened path follows the steps
    simulated: passing particles in decreasing s_i / w_i reaches recall R with
    the least work, a share g(R) of the exhaustive path's steps.  1 / g(R)
    bounds the gain in accepted pairs per second at recall R (1 / f(R) when
    every particle costs the same); the screen's own cost only lowers it.
    Returns f(R), g(R) and their inverses per recall, the share of particles
    with any accepted replicate and the mean pairs per particle."""
    import torch
    dev = torch.device("cuda:0")
    L = eng.L
    G = data.shape[0]
    KS = data.shape[1]
    data_d = torch.tensor(np.asfortranarray(data).ravel(order="F"), dtype=torch.float64, device=dev)
    se_d = torch.tensor(np.asfortranarray(se).ravel(order="F"), dtype=torch.float64, device=dev)
    counts, works = [], []
    for m in sorted(thetas):
        th = thetas[m]
        n = th.numel() // eng.n_params(m)
        st = torch.empty((4 * eng.nb + 3 * eng.J) * n, dtype=torch.float64, device=dev)
        sup = torch.empty(KS * n, dtype=torch.float64, device=dev)
        err = torch.empty(n * G, dtype=torch.float64, device=dev)
        cnt = torch.zeros((G, n), dtype=torch.int32, device=dev)
        stp = torch.empty(n, dtype=torch.int64, device=dev)
        wsum = torch.zeros(n, dtype=torch.float64, device=dev)
        for r in range(reps):
            _lib.check(L.abc_simulate_dev(eng.ctx, m, seed0 + r, firsts[m], n, _ptr(th), _ptr(st),
                                          None, _ptr(sup), _ptr(stp)), eng.ctx, "simulate")
            wsum += stp.to(torch.float64)
            _lib.check(L.abc_errors_dev(eng.ctx, n, KS, _ptr(sup), G, _ptr(data_d), _ptr(se_d),
                                        parts.ctypes.data_as(C.c_void_p), len(parts), 10.0,
                                        _ptr(err)), eng.ctx, "errors")
            cnt += (err.view(G, n) <= eps).to(torch.int32)
        counts.append(cnt)
        works.append(wsum / reps)
        log(f"[ceiling] model {m}: {reps} replicates of {n} particles")
    p = torch.cat(counts, dim=1).to(torch.float64) / reps          # G x n_all
    s_i = p.sum(0).cpu().numpy()
    w_i = torch.cat(works).cpu().numpy()
    out = {"reps": reps}
    out.update(ceiling_shares(s_i, w_i, recalls))
    return out


def ceiling_shares(s_i, w_i, recalls=(0.99, 0.995)):
    """From expected accepted pairs s_i and costs w_i (>= 1 each) per
    particle: the smallest share of the particles whose pairs reach each
    recall R (particles in decreasing s_i), and the share of the work in
    decreasing s_i / w_i, the greedy order of the fractional knapsack: no
    subset of particles reaching R costs less than that share, so its
    inverse bounds the gain in pairs per unit of work."""
    s_i = np.asarray(s_i, np.float64)
    w_i = np.maximum(np.asarray(w_i, np.float64), 1.0)
    if len(s_i) == 0:
        raise ValueError("ceiling_shares: no particles")
    out = {"particles": int(len(s_i)),
           "pairs_per_particle": float(s_i.mean()),
           "share_with_any_acceptance": float((s_i > 0).mean())}
    total = max(float(s_i.sum()), 1e-300)
    for key, order in (("", np.argsort(-s_i, kind="stable")),
                       ("work_", np.argsort(-(s_i / w_i), kind="stable"))):
        cum = np.cumsum(s_i[order]) / total
        wcum = np.cumsum(w_i[order]) / w_i.sum()
        for R in recalls:
            r = min(int(np.searchsorted(cum, R)), len(s_i) - 1)
            if key:
                # the fractional share: the last particle only in part
                need = R * total - (float(np.sum(s_i[order[:r]])))
                part = min(max(need / s_i[order[r]], 0.0), 1.0) if s_i[order[r]] > 0 else 1.0
                f = (float(wcum[r - 1]) if r > 0 else 0.0) + part * w_i[order[r]] / w_i.sum()
            else:
                f = (r + 1) / len(s_i)
            out[f"{key}share_needed_{R}"] = float(f)
            out[f"{key}gain_bound_{R}"] = 1.0 / float(f)
    return out


def _screened(eng, sur, L, data, se, parts, KS, G, mode, eps, k, thr, z, margin, eval_sets,
              eval_steps, eval_first, n_eval, n_all, plain_pairs, t_plain, seed, dev, log,
              rule=None):
    """One screened run over the evaluation particles: pass flags from the z
    rule, or (rule = (score function of (model, features), threshold)) score >=
    threshold."""
    import torch
    nb, J = eng.nb, eng.J
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _lib.check(L.abc_accept_begin(eng.ctx, KS, G, data.ctypes.data_as(C.c_void_p),
                                  se.ctypes.data_as(C.c_void_p),
                                  parts.ctypes.data_as(C.c_void_p), 2, 10.0, mode, eps, k, 0.02),
               eng.ctx, "accept_begin")
    _lib.check(L.abc_accept_set_thresholds(eng.ctx, thr.ctypes.data_as(C.c_void_p)), eng.ctx, "thr")
    # the screen over every model's particles, then one exact simulation of
    # all the particles that pass (one burn-in and one SSA launch)
    t_screen = 0.0
    passed = {}
    ts = time.perf_counter()
    for m in MODELS:
        th, _ = eval_sets[m]
        x = sur.features(m, th)
        if rule is None:
            pas, _ = sur.screen(eng, x, z=z, margin=margin)
        else:
            pas = (rule[0](m, x) >= rule[1]).to(torch.uint8)
        passed[m] = pas
    keeps = {m: torch.nonzero(passed[m], as_tuple=False).flatten() for m in MODELS}
    npass = {m: int(keeps[m].numel()) for m in MODELS}
    t_screen += time.perf_counter() - ts
    ms = [m for m in MODELS if npass[m] > 0]
    bufs = {}
    for m in ms:
        th, _ = eval_sets[m]
        npar = eng.n_params(m)
        th_keep = th.view(npar, n_eval)[:, keeps[m]].contiguous().view(-1)
        pidx = (keeps[m] + eval_first(m)).to(torch.int64)
        st = torch.empty((4 * nb + 3 * J) * npass[m], dtype=torch.float64, device=dev)
        sup = torch.empty(KS * npass[m], dtype=torch.float64, device=dev)
        bufs[m] = (pidx, th_keep, st, sup)
    t_sim0 = time.perf_counter()
    if ms:
        q = len(ms)
        VA = C.c_void_p * q
        _lib.check(L.abc_simulate_multi_idx_dev(
            eng.ctx, q, (C.c_int32 * q)(*ms), seed, (C.c_int64 * q)(*[npass[m] for m in ms]),
            VA(*[bufs[m][0].data_ptr() for m in ms]), VA(*[bufs[m][1].data_ptr() for m in ms]),
            VA(*[bufs[m][2].data_ptr() for m in ms]), None,
            VA(*[bufs[m][3].data_ptr() for m in ms]), None), eng.ctx, "simulate_multi_idx")
    torch.cuda.synchronize()
    t_sim = time.perf_counter() - t_sim0
    for m in ms:
        pidx_h = np.ascontiguousarray(bufs[m][0].cpu().numpy())
        _lib.check(L.abc_accept_add_idx_dev(eng.ctx, npass[m], _ptr(bufs[m][3]),
                                            pidx_h.ctypes.data_as(C.c_void_p)), eng.ctx, "add_idx")
    A_scr = _end(eng, G, G * max(k, 16))
    t_scr = time.perf_counter() - t0
    n_pass_total = sum(npass.values())
    work_all = sum(int(eval_steps[m].sum().item()) for m in MODELS)
    work_pass = sum(int(eval_steps[m][passed[m].bool()].sum().item()) for m in MODELS)
    scr_pairs = {(g, int(i)) for g, (idx, e) in enumerate(A_scr) for i, ee in zip(idx, e)
                 if ee <= thr[g]}
    hit = len(plain_pairs & scr_pairs)
    res = dict(z=z, screened_time_s=t_scr, screen_time_s=t_screen, simulate_time_s=t_sim,
               screened_pass=n_pass_total,
               ssa_fraction_avoided=1.0 - n_pass_total / n_all,
               ssa_work_fraction_avoided=1.0 - work_pass / max(work_all, 1),
               recall=hit / max(len(plain_pairs), 1),
               screened_accepted_per_s=hit / t_scr,
               speedup_accepted_per_s=(hit / t_scr) / (len(plain_pairs) / t_plain))
    log(f"[screened] {('z=%s' % z) if rule is None else ('score >= %.4g' % rule[1])}: "
        f"pass {n_pass_total}/{n_all}, recall {res['recall']:.4f}, "
        f"time {t_scr:.2f}s (screen {t_screen:.3f}s, exact simulation {t_sim:.3f}s) vs plain "
        f"{t_plain:.2f}s")
    return res


def _end(eng, G, cap):
    """Close the acceptance session: per gene (0-based particle indices, errors)."""
    while True:
        counts = np.zeros(G, np.int64)
        idx = np.zeros(max(cap, 1), np.int64)
        eo = np.zeros(max(cap, 1))
        nc = C.c_int64()
        rc = eng.L.abc_accept_end(eng.ctx, counts.ctypes.data_as(C.c_void_p),
                                  idx.ctypes.data_as(C.c_void_p), eo.ctypes.data_as(C.c_void_p),
                                  cap, C.byref(nc))
        if rc == -5:  # capacity: counts are valid, the session stays open
            cap = int(counts.sum())
            continue
        _lib.check(rc, eng.ctx, "accept_end")
        break
    res, o = [], 0
    for g in range(G):
        res.append((idx[o:o + counts[g]] - 1, eo[o:o + counts[g]]))
        o += counts[g]
    return res
