"""Sweep the moment-model restatement over every shipped fixture row.

Reads the reference's data/posterior_estimates/map_sets_<model>.txt and
data/recovered_statistics/<model>/<cond>/*.txt (the output of
scripts/recover_statistics.jl:27-68 at those parameter sets), recomputes every
value with the oracle and prints the distribution of the gaps.  Run in the
development container only (it reads /root/reference):

    python oracle/validate_fixtures.py [--rows N] [--jobs 8]
"""
import argparse
import os
import sys
from multiprocessing import Pool

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.dirname(HERE))
import oracle as O  # noqa: E402
from abc_inference_transcription_amd.design import MODEL_NAMES, REF_COND_LABELS, REF_PULSE, REF_CHASE  # noqa: E402

R = "/root/reference/data/"
Q = ["mean_u", "mean_l", "var_u", "cov_ul", "var_l"]


def one(args):
    m, i, th = args
    cfg = O.moment_cfg([(2 * k + 1) / 10 * 20 for k in range(5)], REF_PULSE, REF_CHASE, 20.0)
    return m, i, O.recover(m, th, cfg)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--rows", type=int, default=10**9)
    ap.add_argument("--jobs", type=int, default=8)
    a = ap.parse_args()
    jobs, fx = [], {}
    for m, mn in enumerate(MODEL_NAMES, 1):
        maps = np.loadtxt(R + f"posterior_estimates/map_sets_{mn}.txt", ndmin=2)
        for c in REF_COND_LABELS:
            for q in Q:
                fx[(m, c, q)] = np.loadtxt(R + f"recovered_statistics/{mn}/{c}/{q}.txt", ndmin=2)
        jobs += [(m, i, maps[i]) for i in range(min(a.rows, len(maps)))]
    with Pool(a.jobs) as p:
        res = p.map(one, jobs, chunksize=4)
    rel_by = {m: [] for m in range(1, 6)}
    abs_small = []
    for m, i, out in res:
        for j, c in enumerate(REF_COND_LABELS):
            for qi, q in enumerate(Q):
                ref = fx[(m, c, q)][i]
                mine = out[j, :, qi]
                big = np.abs(ref) > 1e-3
                rel_by[m] += list(np.abs(mine - ref)[big] / np.abs(ref)[big])
                abs_small += list(np.abs(mine - ref)[~big])
    print(f"rows={len(res)} values={5 * 11 * 5 * len(res)}")
    for m in range(1, 6):
        v = np.array(rel_by[m])
        print(f"{MODEL_NAMES[m-1]:12s} n={len(v):7d} rel gap (|ref|>1e-3): median={np.median(v):.2e} "
              f"p90={np.quantile(v, .9):.2e} p99={np.quantile(v, .99):.2e} p99.9={np.quantile(v, .999):.2e} max={v.max():.2e}")
    s = np.array(abs_small)
    if len(s):
        print(f"abs gap for |ref|<=1e-3: n={len(s)} median={np.median(s):.2e} max={s.max():.2e}")


if __name__ == "__main__":
    main()
