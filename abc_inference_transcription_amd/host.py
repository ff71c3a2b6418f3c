"""Host mirror of the reference's script contract, driven by libabc.so.

The reference has no API: a caller sets the globals ``m``, ``n_trials`` and
``submit`` (wrapper.jl:59-63) and includes three scripts that talk through
tab-delimited text files.  This module keeps those names, their arguments and
the files they write, byte for byte, so that the downstream scripts can read
what it produces:

* :func:`abc_simulation` -- scripts/abc_simulation.jl:88-97 (prior draw,
  simulation, five statistics files plus ``sets_`` and ``progress_``)
* :func:`load_summary_stats` -- the loader wrapper.jl:42 calls (undefined in
  the reference; reads the 14 matrices of data/summary_stats)
* :func:`load_s_data` -- scripts/compute_errors.jl:17-28
* :func:`compute_errors` -- scripts/compute_errors.jl:45-78
* :func:`accepted_particles` -- scripts/accepted_particles.jl:9-32

Floats are printed the way Julia's ``print``/``writedlm`` prints a Float64
(shortest round-trip digits; plain notation when the decimal point falls
within 4 places below and 6 above the first digit, else ``1.5e-5`` style
with no ``+`` and no exponent padding).

All simulation and error work runs in libabc.so through :class:`Engine`.
"""
from __future__ import annotations

import math
import os
from decimal import Decimal

import numpy as np

from .design import MODEL_NAMES, n_ref_params

# ---------------------------------------------------------------- formatting


def julia_repr(x) -> str:
    """Julia's ``print`` of a Float64 or integer."""
    if isinstance(x, (int, np.integer)) and not isinstance(x, bool):
        return str(int(x))
    x = float(x)
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "Inf" if x > 0 else "-Inf"
    if x == 0.0:
        return "-0.0" if math.copysign(1.0, x) < 0 else "0.0"
    sign = "-" if x < 0 else ""
    # shortest round-trip digits (Python's repr and Julia's Ryu both pick the
    # shortest string that reads back to x, nearest to x on ties)
    t = Decimal(repr(abs(x))).normalize().as_tuple()
    digits = "".join(map(str, t.digits))
    nd = len(digits)
    pt = t.exponent + nd  # position of the decimal point relative to the first digit
    plain = -4 < pt <= 6
    if plain and pt >= nd:
        # Julia falls back to exponent form when the integer is not exactly
        # representable at the printed width
        if abs(math.fmod(abs(x) + 0.05, 10.0 ** (pt - nd)) - 0.05) > 0.05:
            plain = False
    if plain:
        if pt <= 0:
            body = "0." + "0" * (-pt) + digits
        elif pt < nd:
            body = digits[:pt] + "." + digits[pt:]
        else:
            body = digits + "0" * (pt - nd) + ".0"
    else:
        body = digits[0] + "." + (digits[1:] or "0") + "e" + str(pt - 1)
    return sign + body


def _rows_text(rows) -> str:
    return "".join("\t".join(julia_repr(v) for v in row) + "\n" for row in rows)


def _append(path: str, text: str):
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    with open(path, "a") as fh:
        fh.write(text)


def readdlm(path: str) -> np.ndarray:
    """Tab/space-delimited numeric matrix (DelimitedFiles.readdlm)."""
    rows = [line.split() for line in open(path) if line.strip()]
    return np.array([[float(v) for v in r] for r in rows], dtype=np.float64)


# ------------------------------------------------------- statistics layout

def split_stats(stats: np.ndarray, n_bins: int, n_conds: int):
    """Split P x (4*nb + 3*J) statistics into the seven reference parts:
    pulse mean, pulse FF, chase mean, chase FF, ratios, mean_corr, corr_mean
    (abc_simulation.jl:47-61, compute_errors.jl:55-57)."""
    nb, J = n_bins, n_conds
    o = np.cumsum([0, nb, nb, nb, nb, J, J, J])
    return [stats[:, o[i]:o[i + 1]] for i in range(7)]


def join_stats(parts) -> np.ndarray:
    return np.asfortranarray(np.hstack([np.asarray(p, np.float64) for p in parts]))


def statistics_text(stats: np.ndarray, n_bins: int, n_conds: int) -> dict:
    """Text appended to each statistics file for a block of particles, in the
    layout of abc_simulation.jl:47-61: s_pulse/s_chase take two lines per
    particle (means, then Fano factors), the three condition statistics one."""
    pm, pf, cm, cf, ra, mc, cmn = split_stats(np.asarray(stats), n_bins, n_conds)
    inter = lambda a, b: [r for i in range(a.shape[0]) for r in (a[i], b[i])]
    return {"s_pulse": _rows_text(inter(pm, pf)), "s_chase": _rows_text(inter(cm, cf)),
            "s_ratios": _rows_text(ra), "s_mean_corr": _rows_text(mc),
            "s_corr_mean": _rows_text(cmn)}


# ----------------------------------------------------------- abc_simulation

def abc_simulation(eng, m: int, n_trials: int, submit: int, root: str = "data/simulations",
                   seed: int = 0, batch: int = 4096, first_particle: int | None = None):
    """scripts/abc_simulation.jl:88-97 for model ``m`` (1..5): ``n_trials``
    particles of job ``submit``.  Particle i (1-based) of the job is global
    particle ``(submit-1)*n_trials + i - 1``, so jobs are disjoint and any
    range can be recomputed bit for bit.  Appends to
    ``<root>/<model>/{progress,sets,s_pulse,s_chase,s_ratios,s_mean_corr,
    s_corr_mean}_<model>_<submit>.txt``.  Returns (theta, stats) of the job."""
    if m < 1 or m > 5:
        raise ValueError("m must be 1..5")
    name = MODEL_NAMES[m - 1]
    base = (submit - 1) * n_trials if first_particle is None else first_particle
    d = os.path.join(root, name)
    fn = lambda stem: os.path.join(d, f"{stem}_{name}_{submit}.txt")
    npar = n_ref_params(m)
    thetas, stats_all = [], []
    for i0 in range(0, n_trials, batch):
        nblk = min(batch, n_trials - i0)
        theta = eng.sample_prior(m, nblk, seed=seed, first=base + i0)
        st = eng.simulate(m, theta, seed=seed, first=base + i0)["stats"]
        _append(fn("progress"), "".join(f"{i0 + i + 1}\n" for i in range(nblk)))
        _append(fn("sets"), _rows_text(theta[:, :npar]))
        for stem, text in statistics_text(st, eng.nb, eng.J).items():
            _append(fn(stem), text)
        thetas.append(theta)
        stats_all.append(st)
    return np.vstack(thetas), np.vstack(stats_all)


# ---------------------------------------------------------- compute_errors

SUMMARY_FILES = ["pulse_mean", "pulse_ff", "pulse_mean_se", "pulse_ff_se", "chase_mean",
                 "chase_ff", "chase_mean_se", "chase_ff_se", "ratio_data", "ratio_se",
                 "mean_corr_data", "mean_corr_se", "corr_mean_data", "corr_mean_se"]


def summary_statistics(eng, u_data, l_data, n_clusters: int, age, experiment, cond_vec,
                       pulse_idx, chase_idx, age_id_distribution, n_boot: int = 100,
                       seed: int = 0, out_dir: str | None = "data/summary_stats"):
    """scripts/data_summary_statistics.jl:183-241: get_summary_stats for every
    gene of the data, in the reference's conventions: u_data, l_data cells x
    genes counts; age the 1-based age cluster of each cell; experiment its
    labelling experiment id and cond_vec the ids of the conditions, in order;
    pulse_idx, chase_idx 1-based cells; age_id_distribution one vector of
    cluster weights per condition.  The 14 statistics (genes x clusters or
    genes x conditions) are appended one row per gene to the files of
    SUMMARY_FILES under out_dir (none when out_dir is None) and returned as a
    dict.  The bootstrap resamples come from Philox streams keyed by seed."""
    u = np.asarray(u_data)
    l = np.asarray(l_data)
    if np.any(u != np.round(u)) or np.any(l != np.round(l)) or u.min(initial=0) < 0 or l.min(initial=0) < 0:
        raise ValueError("u_data and l_data must hold non-negative integer counts")
    cond_vec = list(cond_vec)
    pos = {e: j for j, e in enumerate(cond_vec)}
    cond = np.array([pos.get(e, -1) for e in np.asarray(experiment).tolist()], np.int32)
    w = np.column_stack([np.asarray(v, np.float64) for v in age_id_distribution])
    out = eng.data_stats(u.astype(np.uint32), l.astype(np.uint32), np.asarray(age, np.int64) - 1,
                         cond, n_clusters, len(cond_vec), np.asarray(pulse_idx, np.int64) - 1,
                         np.asarray(chase_idx, np.int64) - 1, w, n_boot=n_boot, seed=seed)
    widths = [n_clusters] * 8 + [len(cond_vec)] * 6
    res, c0 = {}, 0
    for name, wd in zip(SUMMARY_FILES, widths):
        res[name] = np.ascontiguousarray(out[:, c0:c0 + wd])
        c0 += wd
    if out_dir is not None:
        for name in SUMMARY_FILES:
            _append(os.path.join(out_dir, name + ".txt"), _rows_text(res[name]))
    return res


def load_summary_stats(path: str = "data/summary_stats/", ext: str = ".txt") -> dict:
    """The 14 data matrices wrapper.jl:42 unpacks (genes x ages or genes x
    conditions)."""
    return {k: readdlm(os.path.join(path, k + ext)) for k in SUMMARY_FILES}


def data_matrices(S: dict):
    """(data, se) as G x (4*nb + 3*J) in the reference part order."""
    data = join_stats([S["pulse_mean"], S["pulse_ff"], S["chase_mean"], S["chase_ff"],
                       S["ratio_data"], S["mean_corr_data"], S["corr_mean_data"]])
    se = join_stats([S["pulse_mean_se"], S["pulse_ff_se"], S["chase_mean_se"], S["chase_ff_se"],
                     S["ratio_se"], S["mean_corr_se"], S["corr_mean_se"]])
    return data, se


def load_s_data(path: str, model_name: str, ext: str):
    """scripts/compute_errors.jl:17-28: the seven statistics matrices of a
    model (odd rows of s_pulse/s_chase are means, even rows Fano factors)."""
    f = lambda stem: os.path.join(path, model_name, f"{stem}_{model_name}{ext}")
    sp, sc = readdlm(f("s_pulse")), readdlm(f("s_chase"))
    return (sp[0::2], sp[1::2], sc[0::2], sc[1::2], readdlm(f("s_ratios")),
            readdlm(f("s_mean_corr")), readdlm(f("s_corr_mean")))


def compute_errors(eng, model_name: str, summary: dict, path: str = "data/simulations/",
                   ext: str = ".txt", out: str | None = None, batch: int = 65536,
                   clamp: float = 10.0, columns: str | None = None, text: bool = True):
    """scripts/compute_errors.jl:45-78: errors of every simulated particle
    against every gene, clamped at ``clamp``, one row per particle appended to
    ``out`` (default ``data/errors/error_<model>.txt``; the reference writes to
    a fixed cluster path).

    ``columns`` also writes the matrix as a column-major FP64 ``.npy`` file
    (P x G, Fortran order), filled batch by batch: gene g's errors are one
    contiguous column, as in the per-gene columns ``x<g>`` of the JDF file
    that scripts/process_error_files.jl makes from the text file and
    accepted_particles.jl:1-17 reads.  ``text=False`` skips the text file,
    whose size (about 20 bytes per pair) rules it out at a million particles.
    Returns the P x G matrix, or the memory map of ``columns`` when the text
    file is skipped."""
    stats = join_stats(load_s_data(path, model_name, ext))
    data, se = data_matrices(summary)
    out = out or os.path.join("data", "errors", f"error_{model_name}.txt")
    P, G = stats.shape[0], data.shape[0]
    cols = None
    if columns:
        os.makedirs(os.path.dirname(os.path.abspath(columns)), exist_ok=True)
        cols = np.lib.format.open_memmap(columns, mode="w+", dtype=np.float64, shape=(P, G),
                                         fortran_order=True)
    mats = []
    for i0 in range(0, P, batch):
        err = eng.errors(stats[i0:i0 + batch], data, se, clamp=clamp)
        if text:
            _append(out, _rows_text(err))
            mats.append(err)
        if cols is not None:
            cols[i0:i0 + err.shape[0]] = err
    if cols is not None:
        cols.flush()
        if not text:
            return cols
    return np.vstack(mats) if mats else np.zeros((0, G))


# ------------------------------------------------------ accepted_particles

def accepted_particles(eng, summary: dict, stats_by_model: dict, eps: float = 4.8,
                       out_dir: str = "data/posteriors", clamp: float = 10.0):
    """scripts/accepted_particles.jl:9-32: for each gene and each model, the
    1-based indices of particles with error <= eps, ascending by error (ties
    by index), appended as one line to ``particles_<model>.txt``; ``0`` when
    none passes.  ``stats_by_model`` maps a model name to its P x K
    statistics (row i = particle i+1).  The reference loops over
    ``1:lastindex(model_name)`` (a string length, accepted_particles.jl:19);
    this loops over the models given.  Returns {model: per-gene lists}."""
    data, se = data_matrices(summary)
    G = data.shape[0]
    res = {name: eng.accept(st, data, se, clamp=clamp, mode=0, eps=eps)
           for name, st in stats_by_model.items()}
    lines = {name: [] for name in res}
    for g in range(G):
        for name in res:
            idx = res[name][g][0]
            lines[name].append("\t".join(str(int(v)) for v in idx) + "\n" if len(idx) else "0\n")
    for name in res:
        _append(os.path.join(out_dir, f"particles_{name}.txt"), "".join(lines[name]))
    return res


def accepted_from_columns(columns_by_model: dict, eps: float = 4.8,
                          out_dir: str = "data/posteriors"):
    """scripts/accepted_particles.jl:1-32 on error matrices stored by
    ``compute_errors(columns=...)``: for each gene g and each model, the
    1-based indices of the particles whose error in column g is <= eps,
    ascending by error with ties by index (Julia's sortperm is stable),
    appended as one line to ``particles_<model>.txt``; ``0`` when none
    passes.  Each column is read from a memory map, so one gene's errors of
    every model are in memory at a time.  Returns {model: number of accepted
    particles per gene}."""
    mm = {name: np.load(p, mmap_mode="r", allow_pickle=False)
          for name, p in columns_by_model.items()}
    G = {m.shape[1] for m in mm.values()}
    if len(G) != 1:
        raise ValueError("error matrices disagree on the number of genes")
    G = G.pop()
    lines = {name: [] for name in mm}
    counts = {name: np.zeros(G, np.int64) for name in mm}
    for g in range(G):
        for name, m in mm.items():
            col = np.asarray(m[:, g])
            v = np.nonzero(col <= eps)[0]
            counts[name][g] = len(v)
            if len(v):
                idx = v[np.argsort(col[v], kind="stable")] + 1
                lines[name].append("\t".join(map(str, idx.tolist())) + "\n")
            else:
                lines[name].append("0\n")
    for name in mm:
        _append(os.path.join(out_dir, f"particles_{name}.txt"), "".join(lines[name]))
    return counts


# ------------------------------------------------------------ model selection
# The step after acceptance.  Model probabilities come from libabc
# (abc_model_probs); the decisions below are the reference's per-gene rules
# (scripts/model_selection.jl:1-78), on numpy arrays with 1-based gene ids.

def model_probabilities(eng, counts, out_dir: str, prefix: str = "", seed: int = 0):
    """scripts/model_probs.jl:33-63 (and its constant_ / non_constant_
    variants with prefix "c_" / "nc_"): per gene, the model-group
    probabilities and bootstrap bounds from the accepted-particle counts
    (G x M), appended as one row per gene to model_prob.txt, l_bound.txt and
    u_bound.txt.  Returns (prob, l_bound, u_bound)."""
    prob, lb, ub = eng.model_probs(counts, seed=seed)
    os.makedirs(out_dir, exist_ok=True)
    for name, a in (("model_prob", prob), ("l_bound", lb), ("u_bound", ub)):
        _append(os.path.join(out_dir, f"{prefix}{name}.txt"), _rows_text(a))
    return prob, lb, ub


def get_n_particles(posterior):
    """model_selection.jl:71-78: accepted particles per model and gene from the
    per-model posterior lists, where a list [0] means none."""
    out = []
    for p in posterior:
        out.append(np.array([0 if list(x) == [0] else len(x) for x in p], np.int64))
    return out


def get_c_nc_genes(prob, l_b, u_b, gene_vec, n_particles, min_particles: int, threshold: float):
    """model_selection.jl:1-17: constant (1), non-constant (2), undetermined
    (0) or not fitted (3) per gene; returns the four lists of gene ids."""
    sel = []
    for g in gene_vec:
        r = g - 1
        if np.any(prob[r] != 0.0) and n_particles[r] >= min_particles:
            if l_b[r, 0] > threshold:
                sel.append(1)
            elif u_b[r, 0] <= threshold:
                sel.append(2)
            else:
                sel.append(0)
        else:
            sel.append(3)
    gv = np.asarray(gene_vec)
    sel = np.asarray(sel)
    return [gv[sel == 1], gv[sel == 2], gv[sel == 0], gv[sel == 3]]


def cluster_c_genes(prob, l_b, u_b, gene_vec):
    """model_selection.jl:19-31: scaling (1) vs non-scaling (2) among the
    constant genes when one model's lower bound reaches the other's upper."""
    sel = np.zeros(len(gene_vec), np.int64)
    for i, g in enumerate(gene_vec):
        r = g - 1
        if np.any(prob[r] != 0.0):
            if l_b[r, 0] >= u_b[r, 1]:
                sel[i] = 1
            elif l_b[r, 1] >= u_b[r, 0]:
                sel[i] = 2
    gv = np.asarray(gene_vec)
    return [gv[sel == 1], gv[sel == 2], gv[sel == 0]]


def cluster_nc_genes(prob, l_b, u_b, gene_vec):
    """model_selection.jl:34-68: among the non-constant genes, the models
    whose lower bound reaches another's upper bound.  Returns (sel, [one-model
    genes per model, two-model genes per pair, 'maybe' genes per model,
    [undetermined genes]]) with sel[i] a list as in the reference ([m], [m1,
    m2], [-m] or [0])."""
    M = prob.shape[1]
    nc_models = []
    for g in gene_vec:
        r = g - 1
        nc = []
        for m1 in range(M):
            for m2 in [k for k in range(M) if k != m1]:
                if np.any(prob[r, [m1, m2]] != 0.0) and l_b[r, m1] >= u_b[r, m2]:
                    nc.append(m1 + 1)
        nc_models.append(nc)
    sel = []
    for nc in nc_models:
        freq = np.zeros(M, np.int64)
        for m in set(nc):
            freq[m - 1] = nc.count(m)
        if len(nc) == M - 2:
            sel.append([-nc[0]])
        elif len(nc) == M - 1:
            sel.append([nc[0]] if len(set(nc)) == 1 else list(nc))
        elif len(nc) == M:
            sel.append([int(k) + 1 for k in np.nonzero(freq == M - 1)[0]])
        else:
            sel.append([0])
    gv = np.asarray(gene_vec)

    def pick(want):
        return gv[np.array([i for i, s in enumerate(sel) if s == want], dtype=np.int64)]
    one = [pick([m]) for m in range(1, M + 1)]
    two = [pick([m1, m2]) for m1 in range(1, M) for m2 in range(m1 + 1, M + 1)]
    maybe = [pick([-m]) for m in range(1, M + 1)]
    und = [pick([0])]
    return sel, [one, two, maybe, und]
