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


def _fma32(a, b, c):
    """Exact float32 fma, vectorised: the float64 product of two float32 values is exact;
    where the float64 sum is exact too (checked by two-sum) one rounding to float32 is the
    fma's; the other elements are rounded once from exact rationals."""
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    c = c.astype(np.float64)
    p = a * b
    s = p + c
    bp = s - c
    err = (p - bp) + (c - (s - bp))  # two-sum error of p + c
    out = s.astype(np.float32)
    for i in np.flatnonzero(err != 0.0):
        out[i] = np.float32(float(Fraction(float(a[i])) * Fraction(float(b[i])) + Fraction(float(c[i]))))
    return out


def test_unorm_quotient_equals_division_exhaustively():
    # integer samples (U8 / U16 inputs) are value / (2^bits - 1) in every staged tile
    # (common.h, unorm_quotient): every 16-bit value, every full scale
    v = np.arange(65536, dtype=np.float32)
    for bits in range(1, 17):
        m = np.float32(2 ** bits - 1)
        rm = np.float32(1.0) / m
        q0 = v * rm
        r = _fma32(-q0, np.full_like(v, m), v)
        q = _fma32(r, np.full_like(v, rm), q0)
        want = v / m
        bad = np.flatnonzero(q.view(np.uint32) != want.view(np.uint32))
        assert bad.size == 0, (bits, v[bad[:5]], q[bad[:5]], want[bad[:5]])
        if bits >= 8:  # the correction step matters: the plain product misses the division
            assert np.any(q0.view(np.uint32) != want.view(np.uint32))
