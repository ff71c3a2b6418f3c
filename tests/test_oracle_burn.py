"""The burn-in's Poisson means over one interval of the gene path (oracle
burn_accum, which the GPU burn_kernel reproduces bit for bit): given the gene
state, the expected number of molecules alive at the end of an interval of
length D is m(D) = m(0) e^{-kD} + g * int_0^D (c0 + c1 s) e^{-k(D-s)} ds, for
the unspliced molecules (exit rate k = gamma + splicing) and for all of them
(k = gamma).  Checked against adaptive quadrature of that integral, in both
branches of the closed form (the series below kD = 1e-3 and the exponential
form above), with the gene on and off.  The split between unspliced and
spliced never reaches an output (both decay at gamma), so this is the test
that pins the unspliced mean."""
import numpy as np
import pytest
from scipy import integrate


def _exact(m0, g, c0, c1, k, D):
    if not g:
        return m0 * np.exp(-k * D)
    f = lambda s: (c0 + c1 * s) * np.exp(-k * (D - s))
    val, err = integrate.quad(f, 0.0, D, epsabs=0.0, epsrel=1e-13, limit=200)
    return m0 * np.exp(-k * D) + val


CASES = [
    # (m0_u, m0_t, g, c0, c1, ku, kt, D)
    (0.0, 0.0, 1, 5.0, 0.0, 2.0, 0.5, 1.0),
    (3.0, 7.0, 1, 5.0, 0.25, 2.0, 0.5, 3.0),
    (3.0, 7.0, 0, 5.0, 0.25, 2.0, 0.5, 3.0),
    (0.0, 0.0, 1, 40.0, 2.0, 101.0, 1.0, 0.2),        # fast splicing
    (1.0, 1.0, 1, 1e3, 50.0, 1e-2 + 1e-1, 1e-2, 20.0),  # slow decay
    (0.0, 0.0, 1, 8.0, 0.4, 3e-4, 1e-4, 2.0),         # series branch for both
    (2.0, 5.0, 1, 8.0, 0.4, 4.9e-4, 2e-4, 2.0),       # kD just under 1e-3
    (2.0, 5.0, 1, 8.0, 0.4, 5.1e-4, 2e-4, 2.0),       # ku D just over 1e-3
    (0.0, 0.0, 1, 8.0, 0.4, 30.0, 10.0, 1e-6),        # a very short interval
    (9.0, 9.0, 1, 0.0, 3.0, 0.7, 0.3, 5.0),           # size-scaling term only
]


@pytest.mark.parametrize("case", CASES)
def test_burn_step_matches_quadrature(oracle, case):
    mu0, mt0, g, c0, c1, ku, kt, D = case
    lu, lt = oracle.burn_step(mu0, mt0, g, c0, c1, ku, kt, D)
    eu, et = _exact(mu0, g, c0, c1, ku, D), _exact(mt0, g, c0, c1, kt, D)
    assert lu == pytest.approx(eu, rel=1e-11, abs=1e-300), (case, lu, eu)
    assert lt == pytest.approx(et, rel=1e-11, abs=1e-300), (case, lt, et)


def test_burn_step_composes(oracle):
    """Two intervals of the same gene state equal one interval of their total
    length once the linear rate term restarts at the second interval's start
    (the birth rate is c0 + c1 s from each interval's own start)."""
    c0, c1, ku, kt = 6.0, 0.3, 1.7, 0.4
    a, b = 0.8, 1.9
    lu, lt = oracle.burn_step(0.0, 0.0, 1, c0, c1, ku, kt, a)
    lu, lt = oracle.burn_step(lu, lt, 1, c0 + c1 * a, c1, ku, kt, b)
    eu, et = oracle.burn_step(0.0, 0.0, 1, c0, c1, ku, kt, a + b)
    assert lu == pytest.approx(eu, rel=1e-13)
    assert lt == pytest.approx(et, rel=1e-13)


def test_burn_step_rejects_empty_interval(oracle):
    assert oracle.burn_step(1.5, 2.5, 1, 9.0, 1.0, 1.0, 0.5, 0.0) == (1.5, 2.5)
    assert oracle.burn_step(1.5, 2.5, 1, 9.0, 1.0, 1.0, 0.5, -1.0) == (1.5, 2.5)


def _expected_mean(c, k, T, B, ws, scaling):
    """Mean number of molecules alive at ws with the gene on throughout: none at
    the burn-in start -B*T (a cycle start), births at c (1 + a/T) with size
    scaling or c without (a = time since the cycle start), exit rate k, and
    each molecule kept with probability 1/2 at every later cycle start."""
    def advance(m, a):
        f = lambda s: c * ((1.0 + s / T) if scaling else 1.0) * np.exp(-k * (a - s))
        v, _ = integrate.quad(f, 0.0, a, epsabs=0.0, epsrel=1e-13, limit=200)
        return m * np.exp(-k * a) + v
    m = 0.0
    for _ in range(int(np.floor(ws / T)) + int(B)):
        m = 0.5 * advance(m, T)
    return advance(m, ws - np.floor(ws / T) * T)


@pytest.mark.parametrize("model,design_name,log_gamma,log_splice", [
    (2, "reference", -0.5, 0.5),
    (2, "reference", -2.0, 1.5),
    (1, "config", -1.0, 0.0),
    (1, "config", 0.5, -1.0),
])
def test_burn_cells_match_periodic_means(oracle, model, design_name, log_gamma, log_splice):
    """With the gene held on (kon = 1e3, koff = 1e-9 per hour) the Poisson means
    at each cell's window start are deterministic: the unspliced mean leaves at
    gamma + splicing, the total at gamma, both halved at each division, from an
    empty cell B cycles before the sampling cycle."""
    from abc_inference_transcription_amd.design import config_design, reference_design
    d = reference_design(cells=55) if design_name == "reference" else config_design(cells=45)
    od = oracle.OracleDesign(d)
    theta = np.array([3.0, -9.0, 1.5, log_gamma, -0.3, log_splice])
    r = oracle.burn_cells(od, model, 0, 7, theta, d["cells"])
    T = d["cycle"]
    alpha, gamma, splice = 10.0 ** 1.5, 10.0 ** log_gamma, 10.0 ** log_splice
    assert np.all(r["gene"] == 1)
    for i in range(d["cells"]):
        c = alpha * r["capture"][i]
        eu = _expected_mean(c, gamma + splice, T, r["B"][i], r["ws"][i], model != 2)
        et = _expected_mean(c, gamma, T, r["B"][i], r["ws"][i], model != 2)
        assert r["lu"][i] == pytest.approx(eu, rel=1e-9), (i, r["lu"][i], eu)
        assert r["lt"][i] == pytest.approx(et, rel=1e-9), (i, r["lt"][i], et)
