"""Experimental designs for the ABC simulation path.

A design is plain data: cycle length, age bins, labelling conditions, cells per
particle, capture model and age weights.  It stands in for the Julia globals
the reference reads (scripts/abc_simulation.jl:64-85: ``condition_id``,
``cycle``, ``t0``, ``agevec``; scripts/load_process_data.jl:50-83: ``age``,
``pulse_idx``, ``chase_idx``, ``age_id_distribution``; wrapper.jl:36 ``betas``).
"""
from __future__ import annotations

import numpy as np

MODEL_NAMES = ["const", "const_const", "kon", "alpha", "gamma"]

# scripts/abc_simulation.jl:65 -- pulse lengths and chase times (h)
REF_PULSE = [0.25, 0.5, 0.75, 1.0, 2.0, 3.0, 22.0, 22.0, 22.0, 22.0, 22.0]
REF_CHASE = [0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0, 2.0, 4.0, 6.0]
REF_COND_LABELS = ["pulse_15", "pulse_30", "pulse_45", "pulse_60", "pulse_120", "pulse_180",
                   "chase_0", "chase_60", "chase_120", "chase_240", "chase_360"]

# Config workload: 4 pulses and 5 chases after a 22 h pulse, 21 h cycle.
CFG_PULSE = [0.5, 1.0, 2.0, 3.0, 22.0, 22.0, 22.0, 22.0, 22.0]
CFG_CHASE = [0.0, 0.0, 0.0, 0.0, 0.0, 1.0, 2.0, 4.0, 6.0]

AGE_BIN_CENTRES = 0
AGE_UNIFORM = 1
CAPTURE_NONE = 0
CAPTURE_BETA = 1
CAPTURE_EMPIRICAL = 2
PRIOR_REFERENCE = 0
PRIOR_CONFIG = 1


def uniform_weights(n_bins: int, n_conds: int) -> np.ndarray:
    """Age weights normalised per condition (each column sums to one)."""
    return np.full((n_bins, n_conds), 1.0 / n_bins)


def reference_design(cells: int = 550, capture_mode: int = CAPTURE_NONE, betas=None,
                     beta_bin=None, beta_is_chase=None, weights=None,
                     burn_eps: float = 5e-3) -> dict:
    """The reference's design: T = 20 h, five ages at the bin centres
    (load_process_data.jl:52, abc_simulation.jl:72-74), eleven conditions
    (abc_simulation.jl:65); cells are split evenly over (bin, condition)."""
    nb, J = 5, len(REF_PULSE)
    return dict(cycle=20.0, n_bins=nb, n_conds=J, pulse=list(REF_PULSE),
                chase=list(REF_CHASE), is_chase=[0] * 6 + [1] * 5, cells=int(cells),
                age_mode=AGE_BIN_CENTRES, capture_mode=capture_mode, beta_a=2, beta_b=8,
                betas=betas, beta_bin=beta_bin, beta_is_chase=beta_is_chase,
                weights=uniform_weights(nb, J) if weights is None else np.asarray(weights),
                burn_eps=burn_eps, burn_max=16, prior_mode=PRIOR_REFERENCE,
                splice_log10=0.0, split_target=32768.0, split_max=1024)


def reference_shape_design(betas, cells: int = 5445, seed: int = 5) -> dict:
    """The reference's design at its own shape, with its shipped capture
    efficiencies (data/capture_efficiencies.txt, one per cell of the data;
    wrapper.jl:36).  The raw data that tie each of those cells to an age
    cluster and to the pulse or chase experiments are not shipped, so each
    efficiency gets a seeded uniform age bin and pulse/chase group.  5,445
    cells = 99 per (age, condition), the size of the reference data (5,422
    cells) spread over its 55 groups."""
    betas = np.asarray(betas, np.float64)
    rng = np.random.default_rng(seed)
    return reference_design(cells=cells, capture_mode=CAPTURE_EMPIRICAL, betas=betas,
                            beta_bin=rng.integers(0, 5, len(betas)),
                            beta_is_chase=rng.integers(0, 2, len(betas)))


def config_design(cells: int = 2500, prior_mode: int = PRIOR_CONFIG,
                  capture_mode: int = CAPTURE_BETA, burn_eps: float = 5e-3) -> dict:
    """The benchmark workload: T = 21 h, theta ~ U[0,1), five equal bins, nine
    conditions with cells split evenly, Beta(2,8) capture, config priors."""
    nb, J = 5, len(CFG_PULSE)
    return dict(cycle=21.0, n_bins=nb, n_conds=J, pulse=list(CFG_PULSE),
                chase=list(CFG_CHASE), is_chase=[0] * 4 + [1] * 5, cells=int(cells),
                age_mode=AGE_UNIFORM, capture_mode=capture_mode, beta_a=2, beta_b=8,
                betas=None, beta_bin=None, beta_is_chase=None,
                weights=uniform_weights(nb, J), burn_eps=burn_eps, burn_max=16,
                prior_mode=prior_mode, splice_log10=0.0, split_target=32768.0, split_max=1024)


def n_ref_params(model: int) -> int:
    """Columns of the reference parameter sets (abc_simulation.jl:10)."""
    return 5 if model in (1, 2) else 9


def n_params(model: int) -> int:
    """Reference columns plus the log10 splicing rate."""
    return n_ref_params(model) + 1


def stats_layout(n_bins: int, n_conds: int):
    """Part sizes of the reference statistics vector (compute_errors.jl:148,152)."""
    return [n_bins] * 4 + [n_conds] * 3
