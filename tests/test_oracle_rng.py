"""Pin the oracle's random numbers and elementary functions.

Philox4x32-10 against the Random123 known-answer vectors (Salmon et al. 2011,
kat_vectors: philox4x32_10); log/exp against libm within a few ulp; the
128-bit conversion against Python's correctly rounded int -> float.
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
