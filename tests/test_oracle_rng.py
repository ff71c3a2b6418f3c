"""Pin the oracle's random numbers and elementary functions.

Philox4x32-10 against the Random123 known-answer vectors (Salmon et al. 2011,
kat_vectors: philox4x32_10); log/exp against libm within a few ulp; the
single-precision waiting-time log against double-precision log within a few
float ulp, its table against the mathematics it tabulates, its variates
against Exp(1); the 128-bit conversion against Python's correctly rounded
int -> float.
"""
import math
import struct
import random

import numpy as np


KATS = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


def test_philox_known_answers(oracle):
    for ctr, key, want in KATS:
        assert list(oracle.philox(ctr, key)) == want


def test_philox_uniformity(oracle):
    # 4096 blocks of one stream: each output bit is ~fair
    words = np.array([oracle.philox([i, 0, 7, 11], [1, 2]) for i in range(4096)], np.uint64)
    bits = np.unpackbits(words.astype(">u4").view(np.uint8)).reshape(-1, 32)
    frac = bits.mean(0)
    assert np.all(np.abs(frac - 0.5) < 0.05)


def ulps(a, b):
    return abs(a - b) / math.ulp(b) if b != 0 else abs(a)


def test_log_accuracy(oracle):
    # the log is table-driven: 128 subintervals of [0.70703125, 1.4140625)
    # selected by bits 45..51 of x - 0x3fe6a00000000000; sampling every
    # subinterval at several binades pins every table entry
    rng = random.Random(1)
    xs = [rng.uniform(1e-12, 1.0) for _ in range(3000)] + [10 ** rng.uniform(-300, 300) for _ in range(3000)]
    xs += [((w + 0.5) * 2.0 ** -32) for w in (0, 1, 2, 2**31, 2**32 - 1)]
    xs += [5e-324, 2.2250738585072014e-308, 0.5, 2.0, math.sqrt(2), 1 - 2**-53, 1 + 2**-52]
    off = 0x3fe6a00000000000
    for j in range(128):
        for _ in range(12):
            b = off + (j << 45) + rng.randrange(1 << 45)
            for k in (-40, -1, 0, 3):
                xs.append(struct.unpack("<d", struct.pack("<Q", b + (k << 52)))[0])
    worst = max(ulps(oracle.log(x), math.log(x)) for x in xs if x != 1.0)
    assert worst <= 2.0
    assert oracle.log(1.0) == 0.0
    # next to 1 the reduction is exact (1/c = 1): full relative precision
    for d in (2**-52, -2**-53, 1e-10, -1e-10, 2**-9, -2**-9):
        x = 1.0 + d
        assert ulps(oracle.log(x), math.log1p(x - 1.0)) <= 1.0


def _table32(name):
    """A float table of the generated log header, parsed from its hex literals."""
    import os
    import re
    h = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                     "abc_inference_transcription_amd", "csrc", "abc_logtab.h")
    body = re.search(name + r"\[\d+\] = \{(.*?)\};", open(h).read(), re.S).group(1)
    return [float.fromhex(t.strip().rstrip("f")) for t in body.split(",") if t.strip()]


def f32(x):
    return float(np.float32(x))


def test_exp1_32_table_is_the_mathematics():
    # subinterval j of the float bits above 0x3F350000 has midpoint c_j;
    # INV32[j] = float(1/c_j) (1 next to 1) and L32[j] = float(-ln INV32[j])
    inv, L = _table32("abc_log32_inv"), _table32("abc_log32_l")
    assert len(inv) == len(L) == 128
    asf = lambda i: struct.unpack("<f", struct.pack("<I", i))[0]
    one = ((0x3F800000 - 0x3F350000) >> 16) & 127
    for j in range(128):
        c = (asf(0x3F350000 + (j << 16)) + asf(0x3F350000 + ((j + 1) << 16))) / 2
        want = 1.0 if j in (one - 1, one) else f32(1.0 / c)
        assert inv[j] == want, j
        assert abs(L[j] - (-math.log(inv[j]))) <= 0.5 * np.spacing(np.float32(abs(L[j]) or 1e-30)), j


def test_poisson_reciprocal_table_is_exact():
    # the inversion's p_k = p_{k-1} * mu * RN(1/k): entry k must be the
    # correctly rounded quotient 1/k
    rk = _table32("abc_recip_k")
    assert len(rk) == 201 and rk[0] == 0.0
    assert all(rk[k] == 1.0 / k for k in range(1, 201))


def test_exp1_32_accuracy(oracle):
    # against -log of the same single-precision u in double precision; words
    # cover every subinterval of the table in several binades, both ends of the
    # 32-bit range and the exact / rounded boundary of float(w) at 2^24
    rng = random.Random(4)
    words = [0, 1, 2, 3, 2**24 - 1, 2**24, 2**24 + 1, 2**31, 2**31 + 1, 2**32 - 1, 2**32 - 128,
             2**32 - 129, 2**32 - 200, 0x80000000 - 1]
    words += [rng.getrandbits(32) for _ in range(20000)]
    asf = lambda i: struct.unpack("<f", struct.pack("<I", i))[0]
    for j in range(128):
        for e in (0, -1, -5, -20, -31):
            for _ in range(4):
                u = asf(0x3F350000 + (j << 16) + rng.randrange(1 << 16)) * 2.0 ** e
                w = int(u * 2.0 ** 32 - 0.5)
                if 0 <= w < 2**32:
                    words.append(w)
    worst = 0.0
    for w in words:
        u = np.float32(np.float32(np.float32(w) + np.float32(0.5)) * np.float32(2.0 ** -32))
        want = -math.log(float(u))
        got = oracle.exp1_32(w)
        if want == 0.0:
            assert got == 0.0
            continue
        worst = max(worst, abs(got - want) / float(np.spacing(np.float32(want))))
    assert worst <= 3.0, worst
    assert oracle.exp1_32(0) == f32(-math.log(0.5 * 2.0 ** -32))
    # next to u = 1 the reduction is exact: small waiting times keep their
    # relative precision
    for w in (2**32 - 2**9, 2**32 - 2**12, 2**32 - 2**20):
        u = float(np.float32(np.float32(w) + np.float32(0.5)) * np.float32(2.0 ** -32))
        assert abs(oracle.exp1_32(w) - (-math.log1p(u - 1.0))) <= 1e-6 * -math.log1p(u - 1.0)


def test_exp1_32_is_exponential(oracle):
    # Philox words through exp1_32: Kolmogorov-Smirnov against Exp(1)
    from scipy import stats
    ws = [int(x) for i in range(5000) for x in oracle.philox([i, 3, 9, 27], [5, 6])]
    e = np.array([oracle.exp1_32(w) for w in ws])
    assert stats.kstest(e, "expon").pvalue > 1e-3
    assert abs(e.mean() - 1.0) < 4.0 / math.sqrt(len(e))


def test_exp_accuracy(oracle):
    rng = random.Random(2)
    ys = [rng.uniform(-700, 700) for _ in range(4000)] + [rng.uniform(-1, 1) for _ in range(2000)]
    worst = max(ulps(oracle.exp(y), math.exp(y)) for y in ys)
    assert worst <= 4.0
    assert oracle.exp(0.0) == 1.0 and oracle.exp(-800.0) == 0.0
    # exp10 carries the rounding of x*ln10: relative error ~|x| ulp
    zs = [rng.uniform(-6, 6) for _ in range(2000)]
    assert max(abs(oracle.exp10(z) - 10.0 ** z) / 10.0 ** z for z in zs) < 2e-14
    assert oracle.exp10(0.0) == 1.0


def test_exp_scaling_by_one_ldexp(oracle):
    """The device exp scales by 2^k with one ldexp; the oracle in two steps.
    Bit for bit the same over the subnormal results, the normal range and
    small arguments."""
    assert oracle.exp_scaling_mismatches(-745.2, -700.0, 2_000_000, 3) == 0
    assert oracle.exp_scaling_mismatches(-760.0, 710.0, 2_000_000, 5) == 0
    assert oracle.exp_scaling_mismatches(-1e-3, 1e-3, 500_000, 7) == 0
    # the subnormal range is reached: e^-745 ~ 5e-324, e^-708.4 ~ 2.2e-308
    assert 0.0 < oracle.exp(-745.0) < 2.3e-308 and oracle.exp(-708.5) < 2.3e-308


def test_u128_correct_rounding(oracle):
    rng = random.Random(3)
    vals = [0, 1, 2**53 - 1, 2**53, 2**53 + 1, 2**53 + 3, 2**64 - 1, 2**64, 2**127 + 2**74,
            2**127 + 2**75 + 2**74, (2**53 + 1) << 60, ((2**53 + 1) << 60) + 1]
    vals += [rng.getrandbits(rng.randint(1, 128)) for _ in range(5000)]
    for v in vals:
        assert oracle.u128_to_f64(v) == float(v), hex(v)


def test_uniform_conversions(oracle):
    assert oracle.u01_32(0) == 0.5 * 2.0 ** -32
    assert oracle.u01_32(0xFFFFFFFF) < 1.0
    assert oracle.u01_53(0, 0) == 0.0
    assert oracle.u01_53(0xFFFFFFFF, 0xFFFFFFFF) == 1.0 - 2.0 ** -53
