"""The FP64 re-decision of the RAVU hash divides by the constants 12 (five-point
derivative, ravu-r3.hook:64) and pi (angle, ravu-lite-r3.hook:82,85) through a corrected
multiply: q0 = x * RN(1/b), r = fma(-q0, b, x), q = fma(r, RN(1/b), q0)
(mpv-prescalers_amd/csrc/common.h, div_const).  This checks on the CPU that the sequence
returns exactly the correctly rounded quotient x / b that the oracle's division gives,
with fma emulated exactly (rational arithmetic, rounded once by float(Fraction)).
"""
import math
import random
import struct
from fractions import Fraction

import numpy as np
import pytest


def fma(a, b, c):
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def div_const(x, b):
    rb = 1.0 / b
    q0 = x * rb
    return fma(fma(-q0, b, x), rb, q0)


def _random_doubles(rng, n, lo_exp, hi_exp):
    out = []
    for _ in range(n):
        mant = rng.getrandbits(52)
        exp = rng.randrange(1023 + lo_exp, 1023 + hi_exp)
        sign = rng.getrandbits(1)
        out.append(struct.unpack("<d", struct.pack("<Q", (sign << 63) | (exp << 52) | mant))[0])
    return out


def _five_point_numerators(rng, n):
    # numerators of the derivative from FP32 window samples in [0, 1] (full FP32 and 10-bit)
    nums = []
    for i in range(n):
        if i % 2:
            v = [float(np.float32(rng.randrange(1024) / 1023.0)) for _ in range(4)]
        else:
            v = [float(np.float32(rng.random())) for _ in range(4)]
        nums.append(-v[0] + 8.0 * v[1] - 8.0 * v[2] + v[3])
    return nums


@pytest.mark.parametrize("b", [12.0, math.pi])
def test_corrected_multiply_equals_division(b):
    rng = random.Random(20240612)
    xs = _random_doubles(rng, 6000, -60, 10) + _five_point_numerators(rng, 6000)
    xs += [1.0, -1.0, 3.0, 12.0, 0.5, math.pi, 24.0 * math.pi, 2.0 ** -30, 7.0 / 1023.0]
    bad = [(x, x / b, div_const(x, b)) for x in xs if div_const(x, b) != x / b]
    assert not bad, bad[:5]


def test_oracle_rounding_is_what_a_plain_multiply_misses():
    # the plain product x * RN(1/12) differs from x / 12 on some inputs, so the correction
    # step is needed for bucket decisions to follow the oracle exactly
    rng = random.Random(7)
    xs = _random_doubles(rng, 4000, -10, 10)
    assert any(x * (1.0 / 12.0) != x / 12.0 for x in xs)
    assert all(div_const(x, 12.0) == x / 12.0 for x in xs)
