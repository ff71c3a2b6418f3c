"""The drop-in at the reference's own shape: T = 20 h, ages at the five bin
centres, eleven conditions, capture from the shipped efficiencies
(data/capture_efficiencies.txt), the 53 statistics of abc_sim against all
3,419 genes of data/summary_stats, and the fixed epsilon 4.8 of
scripts/accepted_particles.jl:10 (wrapper.jl:59-78, compute_errors.jl:51-64).
Per model, the accepted lists of every gene must equal the oracle's FP64
acceptance of the same statistics (indices, order and error bits); a slice of
particles is replayed by the oracle's simulation bit for bit."""
import os

import numpy as np
import pytest

from abc_inference_transcription_amd.design import reference_shape_design
from abc_inference_transcription_amd.host import data_matrices

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MODELS = (1, 2, 3, 4, 5)
PER_MODEL, EPS = 400, 4.8


def same_bits(a, b):
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.fixture(scope="module")
def refshape(golden_capture):
    from abc_inference_transcription_amd.engine import Engine
    design = reference_shape_design(golden_capture)
    eng = Engine(design)
    S = dict(np.load(os.path.join(GOLDEN, "summary_stats_all.npz")))
    data, se = data_matrices(S)
    assert data.shape == (3419, 53)
    yield eng, design, np.asfortranarray(data), np.asfortranarray(se)
    eng.close()


def test_reference_shape_acceptance_matches_oracle(oracle, refshape):
    eng, design, data, se = refshape
    n_acc = 0
    for m in MODELS:
        print(f"[refshape] model {m}: {PER_MODEL} particles x {design['cells']} cells, "
              f"3419 genes", flush=True)
        th = eng.sample_prior(m, PER_MODEL, first=1000)
        st = eng.simulate(m, th, first=1000)["stats"]
        got = eng.accept(st, data, se, mode=0, eps=EPS)
        err = oracle.errors(st, data, se, eng.parts)
        for g in range(data.shape[0]):
            wi, we = oracle.accept_gene(err[:, g], EPS)
            gi, ge = got[g]
            assert np.array_equal(gi, wi), (m, g)
            assert same_bits(ge, we), (m, g)
            n_acc += len(wi)
    print(f"[refshape] {n_acc} accepted (particle, gene) pairs", flush=True)


def test_reference_shape_oracle_replay(oracle, refshape):
    eng, design, _, _ = refshape
    od = oracle.OracleDesign(design)
    for m in (2, 4):
        th = eng.sample_prior(m, 40, first=1000)
        got = eng.simulate(m, th, first=1000, superset=True)
        # the two particles with the least work (the reference prior spans six
        # decades of rates, so the heaviest would hold the CPU for minutes)
        for i in np.argsort(got["steps"])[:2]:
            print(f"[refshape] oracle replay: model {m} particle {1000 + i}, "
                  f"{int(got['steps'][i])} draws", flush=True)
            want = oracle.simulate(od, m, 0, 1000 + int(i), th[i:i + 1], cells=False)
            assert same_bits(got["stats"][i:i + 1], want["stats_ref"]), (m, i)
            assert same_bits(got["superset"][i:i + 1], want["superset"]), (m, i)
            assert int(got["steps"][i]) == int(want["steps"][0]), (m, i)
