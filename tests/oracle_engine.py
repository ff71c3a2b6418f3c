"""An Engine-shaped object over the CPU oracle, for testing the host mirror
(abc_inference_transcription_amd/host.py) without a GPU and for producing the
expected files the GPU run must reproduce byte for byte."""
import numpy as np


class OracleEngine:
    def __init__(self, O, design):
        self.O = O
        self.od = O.OracleDesign(design)
        self.nb, self.J, self.C = int(design["n_bins"]), int(design["n_conds"]), int(design["cells"])
        self.parts = [self.nb] * 4 + [self.J] * 3

    def n_params(self, model):
        return self.O.n_params(model)

    def sample_prior(self, model, n, seed=0, first=0):
        return self.O.sample_prior(self.od, model, seed, first, n)

    def simulate(self, model, theta, seed=0, first=0, **_):
        r = self.O.simulate(self.od, model, seed, first, theta, cells=False)
        return {"stats": r["stats_ref"], "steps": r["steps"]}

    def errors(self, stats, data, se, parts=None, clamp=10.0):
        return self.O.errors(stats, data, se, self.parts if parts is None else parts, clamp)

    def accept(self, stats, data, se, parts=None, clamp=10.0, mode=0, eps=4.8, **_):
        assert mode == 0
        err = self.errors(stats, data, se, parts, clamp)
        return [self.O.accept_gene(err[:, g], eps) for g in range(err.shape[1])]

    def data_stats(self, u, l, age, cond, n_clusters, n_conds, pulse, chase, age_id_dist,
                   n_boot=100, seed=0):
        return self.O.data_stats(u, l, age, cond, n_clusters, n_conds, pulse, chase, age_id_dist,
                                 n_boot=n_boot, seed=seed)
