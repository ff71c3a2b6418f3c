"""Engine: one libabc context bound to a design, with numpy in/out.

Every computation runs in libabc.so on the GPU; this module only marshals
arrays (column-major, as the Julia host hands them over).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .design import stats_layout


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


class MomentCfg:
    """Configuration of the moment model (the reference's run_part_sim /
    abc_sim arguments, recover_statistics.jl:29-44, abc_simulation.jl:64-76).
    Defaults are the reference's: 20 h cycle, 5 rate bins, t0 = -3 cycles,
    the 1 % transient rule; iv is the transient start (recover_statistics.jl
    starts from E g = 1/2, abc_simulation.jl from E u = 1/2).  method 1 (the
    default) solves the stiff gene block in closed form on each smooth piece
    and integrates the other five moments; method 0 integrates all nine."""

    def __init__(self, ages, pulse, chase, is_chase=None, cycle=20.0, n_steps=5, t0=None,
                 iv=None, transient_mode=0, rtol=1e-10, atol=1e-12, method=1):
        self.ages = np.ascontiguousarray(ages, np.float64)
        self.pulse = np.ascontiguousarray(pulse, np.float64)
        self.chase = np.ascontiguousarray(chase, np.float64)
        self.is_chase = None if is_chase is None else np.ascontiguousarray(is_chase, np.int32)
        iv = [0.5] + [0.0] * 8 if iv is None else list(iv)
        self.s = _lib.AbcMomentCfg(
            float(cycle), int(n_steps), len(self.ages), _p(self.ages), len(self.pulse),
            _p(self.pulse), _p(self.chase), _p(self.is_chase),
            -3.0 * cycle if t0 is None else float(t0), (C.c_double * 9)(*iv),
            int(transient_mode), float(rtol), float(atol), int(method))

    @property
    def n_ages(self):
        return len(self.ages)

    @property
    def n_conds(self):
        return len(self.pulse)


def capture_moments(betas, bins, is_chase, n_ages):
    """Per (age, pulse/chase) group: mean, mean of squares and n-1 variance of
    the cells' capture efficiencies (model.jl:226-234), as the flat array
    abc_moment_sim expects: [(age * 2 + group) * 3 + k]."""
    b = np.asarray(betas, np.float64)
    out = np.zeros((n_ages, 2, 3))
    for a in range(n_ages):
        for g in range(2):
            x = b[(np.asarray(bins) == a) & (np.asarray(is_chase) == g)]
            if len(x) == 0:
                continue
            out[a, g] = [x.mean(), (x * x).mean(), x.var(ddof=1) if len(x) > 1 else np.nan]
    return out.ravel()


class Engine:
    def __init__(self, design: dict, device: int = 0):
        self.L = _lib.load()
        self.design = design
        d = design
        nb, J = int(d["n_bins"]), int(d["n_conds"])
        self.nb, self.J, self.C = nb, J, int(d["cells"])
        self._keep = []

        def arr(x, dt):
            a = np.ascontiguousarray(np.asarray(x, dtype=dt).reshape(-1, order="F"))
            self._keep.append(a)
            return a

        pulse, chase = arr(d["pulse"], np.float64), arr(d["chase"], np.float64)
        is_chase = arr(d["is_chase"], np.int32)
        w = arr(np.asarray(d["weights"], np.float64).reshape((nb, J), order="F"), np.float64)
        betas = bbin = bch = None
        nbeta = 0
        if d.get("betas") is not None:
            betas = arr(d["betas"], np.float64)
            bbin = arr(d["beta_bin"], np.int32)
            bch = arr(d["beta_is_chase"], np.int32)
            nbeta = len(betas)
        s = _lib.AbcDesign(
            float(d["cycle"]), nb, J, pulse.ctypes.data, chase.ctypes.data, is_chase.ctypes.data,
            self.C, int(d["age_mode"]), int(d["capture_mode"]), int(d.get("beta_a", 2)),
            int(d.get("beta_b", 8)), None if betas is None else betas.ctypes.data,
            None if bbin is None else bbin.ctypes.data, None if bch is None else bch.ctypes.data,
            nbeta, w.ctypes.data, float(d.get("burn_eps", 5e-3)), int(d.get("burn_max", 16)),
            int(d["prior_mode"]), float(d.get("splice_log10", 0.0)),
            float(d.get("split_target", 32768.0)), int(d.get("split_max", 1024)))
        ctx = C.c_void_p()
        rc = self.L.abc_create(C.byref(ctx), device, C.byref(s))
        if rc != 0:
            raise RuntimeError(f"abc_create failed ({rc}): {_lib.ERRORS.get(rc, '?')}")
        self.ctx = ctx

    def close(self):
        if getattr(self, "ctx", None):
            self.L.abc_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        _lib.check(rc, self.ctx, what)

    # ------------------------------------------------------------------ API
    def n_params(self, model: int) -> int:
        return self.L.abc_n_params(self.ctx, model)

    @property
    def n_stats(self) -> int:
        return self.L.abc_n_stats(self.ctx)

    @property
    def parts(self):
        return stats_layout(self.nb, self.J)

    def sample_prior(self, model: int, n: int, seed: int = 0, first: int = 0) -> np.ndarray:
        th = np.zeros((n, self.n_params(model)), order="F")
        self._chk(self.L.abc_sample_prior(self.ctx, model, seed, first, n, _p(th)), "sample_prior")
        return th

    def simulate(self, model: int, theta, seed: int = 0, first: int = 0, cells=False,
                 moments=False, superset=False) -> dict:
        theta = np.asfortranarray(theta, dtype=np.float64)
        n = theta.shape[0]
        if theta.shape[1] != self.n_params(model):
            raise ValueError(f"theta needs {self.n_params(model)} columns for model {model}")
        NBJ = self.nb * self.J
        out = {"stats": np.zeros((n, self.n_stats), order="F"),
               "steps": np.zeros(n, np.int64)}
        if moments:
            out["moments"] = np.zeros((n, NBJ * 6), order="F")
        if superset:
            out["superset"] = np.zeros((n, NBJ * 2), order="F")
        if cells:
            out["cell_u"] = np.zeros((n, self.C), np.uint32)
            out["cell_l"] = np.zeros((n, self.C), np.uint32)
            out["cell_bin"] = np.zeros((n, self.C), np.uint8)
        rc = self.L.abc_simulate(self.ctx, model, seed, first, n, _p(theta), _p(out["stats"]),
                                 _p(out.get("moments")), _p(out.get("superset")),
                                 _p(out["steps"]), _p(out.get("cell_u")), _p(out.get("cell_l")),
                                 _p(out.get("cell_bin")))
        self._chk(rc, "simulate")
        return out

    def errors(self, stats, data, se, parts=None, clamp: float = 10.0) -> np.ndarray:
        stats = np.asfortranarray(stats, np.float64)
        data = np.asfortranarray(data, np.float64)
        se = np.asfortranarray(se, np.float64)
        n, K = stats.shape
        G = data.shape[0]
        parts = np.ascontiguousarray(self.parts if parts is None else parts, np.int32)
        err = np.zeros((n, G), order="F")
        rc = self.L.abc_errors(self.ctx, n, K, _p(stats), G, _p(data), _p(se), _p(parts),
                               len(parts), clamp, _p(err))
        self._chk(rc, "errors")
        return err

    def accept(self, stats, data, se, parts=None, clamp: float = 10.0, mode: int = 0,
               eps: float = 4.8, k: int = 1000, first_index: int = 0, cap: int | None = None):
        """Per gene: (1-based particle indices, FP64 errors), sorted by (err, index)."""
        stats = np.asfortranarray(stats, np.float64)
        data = np.asfortranarray(data, np.float64)
        se = np.asfortranarray(se, np.float64)
        n, K = stats.shape
        G = data.shape[0]
        parts = np.ascontiguousarray(self.parts if parts is None else parts, np.int32)
        if cap is None:
            cap = G * (k if mode == 1 else min(n, 4096)) + 16
        while True:
            counts = np.zeros(G, np.int64)
            idx = np.zeros(cap, np.int64)
            eo = np.zeros(cap)
            rc = self.L.abc_accept(self.ctx, n, K, _p(stats), G, _p(data), _p(se), _p(parts),
                                   len(parts), clamp, mode, eps, k, first_index, _p(counts),
                                   _p(idx), _p(eo), cap)
            if rc == -5:
                cap = int(counts.sum()) + 16
                continue
            self._chk(rc, "accept")
            break
        out, o = [], 0
        for g in range(G):
            out.append((idx[o:o + counts[g]].copy(), eo[o:o + counts[g]].copy()))
            o += counts[g]
        return out

    def screen_errors(self, stats, data, se, parts=None, clamp: float = 10.0):
        """Approximate errors of the split-BF16 GEMM and their proven bounds."""
        stats = np.asfortranarray(stats, np.float64)
        data = np.asfortranarray(data, np.float64)
        se = np.asfortranarray(se, np.float64)
        n, K = stats.shape
        G = data.shape[0]
        parts = np.ascontiguousarray(self.parts if parts is None else parts, np.int32)
        approx = np.zeros((n, G), np.float32, order="F")
        bound = np.zeros((n, G), np.float32, order="F")
        rc = self.L.abc_screen_errors(self.ctx, n, K, _p(stats), G, _p(data), _p(se), _p(parts),
                                      len(parts), clamp, _p(approx), _p(bound))
        self._chk(rc, "screen_errors")
        return approx, bound

    def accept_stream(self, batches, data, se, parts=None, clamp: float = 10.0, mode: int = 0,
                      eps: float = 4.8, k: int = 1000, margin_rel: float = 0.02):
        """Streaming acceptance over an iterable of (stats, first_index) batches.
        Returns (per-gene lists, number of FP64 rescorings)."""
        data = np.asfortranarray(data, np.float64)
        se = np.asfortranarray(se, np.float64)
        G, K = data.shape
        parts = np.ascontiguousarray(self.parts if parts is None else parts, np.int32)
        self._chk(self.L.abc_accept_begin(self.ctx, K, G, _p(data), _p(se), _p(parts), len(parts),
                                          clamp, mode, eps, k, margin_rel), "accept_begin")
        for stats, first in batches:
            stats = np.asfortranarray(stats, np.float64)
            self._chk(self.L.abc_accept_add(self.ctx, stats.shape[0], _p(stats), first),
                      "accept_add")
        cap = G * max(k, 1) + 16 if mode == 1 else 1 << 20
        counts = np.zeros(G, np.int64)
        idx = np.zeros(cap, np.int64)
        eo = np.zeros(cap)
        ncand = C.c_int64()
        self._chk(self.L.abc_accept_end(self.ctx, _p(counts), _p(idx), _p(eo), cap,
                                        C.byref(ncand)), "accept_end")
        out, o = [], 0
        for g in range(G):
            out.append((idx[o:o + counts[g]].copy(), eo[o:o + counts[g]].copy()))
            o += counts[g]
        return out, ncand.value

    def last_timing(self):
        ms = np.zeros(4)
        self._chk(self.L.abc_last_timing(self.ctx, _p(ms)), "last_timing")
        return ms

    def bind_stream(self, torch_stream):
        """Launch every later call on this (non-null) torch stream."""
        self._chk(self.L.abc_set_stream(self.ctx, C.c_void_p(torch_stream.cuda_stream)),
                  "set_stream")
        self._stream = torch_stream

    def set_ssa_blocks(self, blocks: int):
        self._chk(self.L.abc_set_ssa_blocks(self.ctx, blocks), "set_ssa_blocks")

    def ssa_blocks(self) -> int:
        return self.L.abc_get_ssa_blocks(self.ctx)

    def set_ssa_claim(self, claim: int):
        self._chk(self.L.abc_set_ssa_claim(self.ctx, claim), "set_ssa_claim")

    def ssa_claim(self) -> int:
        return self.L.abc_get_ssa_claim(self.ctx)

    def autotune(self, models=(1, 2, 3, 4, 5), n: int = 256, seed: int = 0, first: int = 0):
        """Pick the fastest persistent-SSA grid for this workload (results do not
        depend on it).  Returns (workgroups, ms of the SSA over the sample)."""
        ms = np.array(list(models), np.int32)
        best, t = C.c_int32(), C.c_double()
        self._chk(self.L.abc_autotune_ssa(self.ctx, len(ms), _p(ms), seed, first, n,
                                          C.byref(best), C.byref(t)), "autotune_ssa")
        return best.value, t.value

    def capture(self, fn, graph=None):
        """Record the *_dev calls ``fn()`` makes on this engine's stream into a
        HIP graph (a new one, or an in-place update of ``graph``)."""
        self._chk(self.L.abc_graph_begin(self.ctx), "graph_begin")
        try:
            fn()
        finally:
            h = C.c_void_p(graph)
            rc = self.L.abc_graph_end(self.ctx, C.byref(h))
        self._chk(rc, "graph_end")
        return h.value

    def replay(self, graph):
        self._chk(self.L.abc_graph_launch(self.ctx, C.c_void_p(graph)), "graph_launch")

    def recover(self, model: int, theta, cfg: MomentCfg, steps=False):
        """The reference's moment model at the given parameter rows (theta:
        n x reference columns, log10; extra columns are ignored) without
        downsampling: (n, n_conds, n_ages, 5) of mean_u, mean_l, var_u, cov_ul,
        var_l (recover_statistics.jl:2-11)."""
        th = np.asfortranarray(theta, np.float64)
        n = th.shape[0]
        out = np.zeros((n, cfg.n_conds * cfg.n_ages * 5), order="F")
        st = np.zeros(n, np.int64) if steps else None
        self._chk(self.L.abc_recover(self.ctx, model, n, _p(th), n, C.byref(cfg.s), _p(out),
                                     _p(st)), "recover")
        out = out.reshape(n, cfg.n_conds, cfg.n_ages, 5)
        return (out, st) if steps else out

    def moment_sim(self, model: int, theta, cfg: MomentCfg, weights, capture=None,
                   pulse_cond: int = 5, chase_cond: int = 6, steps=False):
        """abc_sim on the moment model (abc_simulation.jl:13-62): returns the
        statistics (n x (4 n_ages + 3 n_conds), reference order) and the
        downsampled moments (n, n_conds, n_ages, 5)."""
        th = np.asfortranarray(theta, np.float64)
        n = th.shape[0]
        K = 4 * cfg.n_ages + 3 * cfg.n_conds
        w = np.ascontiguousarray(np.asarray(weights, np.float64).reshape(-1, order="F"))
        cap = None if capture is None else np.ascontiguousarray(capture, np.float64)
        stats = np.zeros((n, K), order="F")
        mom = np.zeros((n, cfg.n_conds * cfg.n_ages * 5), order="F")
        st = np.zeros(n, np.int64) if steps else None
        self._chk(self.L.abc_moment_sim(self.ctx, model, n, _p(th), n, C.byref(cfg.s), _p(cap),
                                        _p(w), pulse_cond, chase_cond, _p(stats), _p(mom),
                                        _p(st)), "moment_sim")
        mom = mom.reshape(n, cfg.n_conds, cfg.n_ages, 5)
        return (stats, mom, st) if steps else (stats, mom)

    def model_probs(self, counts, seed: int = 0, n_boot: int = 100, alpha: float = 0.95):
        """get_model_probs for every gene (scripts/model_probs.jl:1-63):
        counts G x M accepted particles per model group -> (prob, l_bound,
        u_bound), each G x M."""
        cn = np.asfortranarray(counts, np.int64)
        G, M = cn.shape
        prob, lb, ub = (np.zeros((G, M), order="F") for _ in range(3))
        self._chk(self.L.abc_model_probs(self.ctx, G, M, _p(cn), seed, n_boot, alpha, _p(prob),
                                         _p(lb), _p(ub)), "model_probs")
        return prob, lb, ub

    def data_stats(self, u, l, age, cond, n_clusters, n_conds, pulse, chase, age_id_dist,
                   n_boot: int = 100, seed: int = 0):
        """get_summary_stats for every gene (scripts/data_summary_statistics.jl:
        1-196): u, l (cells x genes) integer counts, age 0..n_clusters-1,
        cond 0..n_conds-1 or -1, pulse / chase 0-based cell lists,
        age_id_dist (n_clusters x n_conds).  Returns the genes x (8 n_clusters
        + 6 n_conds) matrix whose column blocks are the 14 summary files."""
        u = np.asfortranarray(u, np.uint32)
        l = np.asfortranarray(l, np.uint32)
        n, G = u.shape
        if l.shape != u.shape:
            raise ValueError("u and l must have the same shape")
        age = np.ascontiguousarray(age, np.int32)
        cond = np.ascontiguousarray(cond, np.int32)
        pulse = np.ascontiguousarray(pulse, np.int32)
        chase = np.ascontiguousarray(chase, np.int32)
        w = np.asfortranarray(age_id_dist, np.float64)
        if len(age) != n or len(cond) != n or w.shape != (n_clusters, n_conds):
            raise ValueError("age, cond and age_id_dist do not match the cells and groups")
        out = np.zeros((G, 8 * n_clusters + 6 * n_conds), order="F")
        self._chk(self.L.abc_data_stats(self.ctx, G, n, _p(u), _p(l), n_clusters, _p(age), n_conds,
                                        _p(cond), _p(pulse), len(pulse), _p(chase), len(chase),
                                        _p(w), n_boot, seed, _p(out)), "data_stats")
        return out

    def drop_graph(self, graph):
        if graph:
            self.L.abc_graph_destroy(C.c_void_p(graph))
