"""Extract small fixture slices from the reference's shipped data files.

Run in the development container (reads /root/reference/data):
    python tests/golden/make_golden.py
Writes tests/golden/*.npz -- inputs and expected outputs only (data, no code):
  recovered.npz : map_sets rows 0..R-1 per model and the matching
                  recovered_statistics rows (scripts/recover_statistics.jl:27-68
                  output; 11 conditions x 5 ages x 5 moments).
  summary_stats.npz : the first G genes of the 14 data/summary_stats matrices.
  summary_stats_all.npz : all 3,419 genes of the 14 matrices (the reference's
                  own error-sweep input at its full size).
  capture.npz   : data/capture_efficiencies.txt (all 5,422 cells).
  julia_floats.txt : float tokens exactly as Julia's writedlm printed them in
                  the reference's data files (every exponent-form token up to
                  4,000, 4,000 plain ones, every one >= 1e5), one per line.
"""
import os

import numpy as np

R = "/root/reference/data/"
OUT = os.path.dirname(os.path.abspath(__file__))
MODELS = ["const", "const_const", "kon", "alpha", "gamma"]
CONDS = ["pulse_15", "pulse_30", "pulse_45", "pulse_60", "pulse_120", "pulse_180",
         "chase_0", "chase_60", "chase_120", "chase_240", "chase_360"]
Q = ["mean_u", "mean_l", "var_u", "cov_ul", "var_l"]
SS = ["pulse_mean", "pulse_ff", "pulse_mean_se", "pulse_ff_se", "chase_mean", "chase_ff",
      "chase_mean_se", "chase_ff_se", "ratio_data", "ratio_se", "mean_corr_data",
      "mean_corr_se", "corr_mean_data", "corr_mean_se"]


def main(rows=12, genes=20):
    rec = {}
    for mn in MODELS:
        maps = np.loadtxt(R + f"posterior_estimates/map_sets_{mn}.txt", ndmin=2)
        r = min(rows, len(maps))
        rec[f"{mn}__theta"] = maps[:r]
        stats = np.zeros((r, len(CONDS), 5, 5))
        for j, c in enumerate(CONDS):
            for qi, q in enumerate(Q):
                stats[:, j, :, qi] = np.loadtxt(R + f"recovered_statistics/{mn}/{c}/{q}.txt", ndmin=2)[:r]
        rec[f"{mn}__stats"] = stats
    np.savez_compressed(os.path.join(OUT, "recovered.npz"), **rec)
    ss = {k: np.loadtxt(R + f"summary_stats/{k}.txt", ndmin=2)[:genes] for k in SS}
    np.savez_compressed(os.path.join(OUT, "summary_stats.npz"), **ss)
    full = {k: np.loadtxt(R + f"summary_stats/{k}.txt", ndmin=2) for k in SS}
    np.savez_compressed(os.path.join(OUT, "summary_stats_all.npz"), **full)
    np.savez_compressed(os.path.join(OUT, "capture.npz"),
                        betas=np.loadtxt(R + "capture_efficiencies.txt"))


def julia_floats(n_exp=4000, n_plain=4000):
    import glob
    files = sorted(glob.glob(R + "summary_stats/*.txt") +
                   glob.glob(R + "recovered_statistics/*/*/*.txt") +
                   glob.glob(R + "posterior_estimates/*.txt"))
    toks = set()
    for f in files:
        for line in open(f):
            toks.update(line.split())
    toks = sorted(toks)
    rng = np.random.default_rng(0)
    expo = [t for t in toks if "e" in t]
    plain = [t for t in toks if "e" not in t]
    big = [t for t in plain if abs(float(t)) >= 1e5]
    pick = lambda L, k: [L[i] for i in sorted(rng.choice(len(L), min(k, len(L)), replace=False))]
    out = sorted(set(pick(expo, n_exp) + pick(plain, n_plain) + big))
    with open(os.path.join(OUT, "julia_floats.txt"), "w") as fh:
        fh.write("\n".join(out) + "\n")


if __name__ == "__main__":
    main()
    julia_floats()
