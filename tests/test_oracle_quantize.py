"""The oracle's quantiser against exact arithmetic.

src/cuSZp_kernels_1D_f32.cu:3-20 computes t = RN_f32(x*r), y = RN_f32(t+0.5),
q = cvt.rzi.s32(y) - (t >= -0.5 ? 0 : 1). Here the same quantity is derived
independently: the product of two f32 values and the sum t+0.5 are exact in
f64, so one rounding to f32 per step reproduces RN_f32; saturation and NaN
follow the PTX cvt rules. For |t| < 2^22 the result also has the closed form
"round half away from zero, except -0.5 -> 0".
"""
import math

import numpy as np
import pytest

I32_MIN, I32_MAX = -(2**31), 2**31 - 1


def rn32(v: float) -> float:
    return float(np.float32(v))


def sat_trunc(y: float) -> int:
    if math.isnan(y):
        return 0
    if y >= 2**31:
        return I32_MAX
    if y <= -(2**31) - 1:
        return I32_MIN
    return int(y)   # toward zero


def wrap(v: int) -> int:
    return ((v + 2**31) % 2**32) - 2**31


def exact_q32(x: np.float32, r: np.float32) -> int:
    t = rn32(float(x) * float(r))          # exact f64 product, one rounding
    y = rn32(t + 0.5)                      # exact f64 sum, one rounding
    s = 0 if t >= -0.5 else 1
    return wrap(sat_trunc(y) - s)


def exact_q64(x: float, r: float) -> int:
    from fractions import Fraction

    def rn64(fr):
        return float(fr)   # Fraction -> float rounds to nearest-even
    if math.isnan(x) or math.isinf(x):
        t = x * r
    else:
        t = rn64(Fraction(x) * Fraction(r))
    y = t + 0.5 if (math.isnan(t) or math.isinf(t)) else rn64(Fraction(t) + Fraction(1, 2))
    s = 0 if t >= -0.5 else 1
    return wrap(sat_trunc(y) - s)


def closed_form(t: float) -> int:
    if t == -0.5:
        return 0
    return int(math.copysign(math.floor(abs(t) + 0.5), t))


def test_ties_and_small_values(oracle):
    for t in [0.0, -0.0, 0.49, 0.5, -0.5, 0.51, -0.51, 1.5, -1.5, 2.5, -2.5, 3.4999, -3.5001, 1e-30, -1e-30]:
        assert oracle.quantize_f32(t, 1.0) == closed_form(t), t
        assert oracle.quantize_f64(t, 1.0) == closed_form(t), t


def test_random_f32_against_exact(oracle):
    rng = np.random.default_rng(0)
    for _ in range(20000):
        e = rng.uniform(-12, 12)
        x = np.float32(rng.choice([-1, 1]) * 10**e)
        r = np.float32(0.5) / np.float32(10 ** rng.uniform(-8, 2))
        assert oracle.quantize_f32(float(x), float(r)) == exact_q32(x, r)


def test_random_f32_closed_form(oracle):
    rng = np.random.default_rng(1)
    for _ in range(20000):
        t = np.float32(rng.uniform(-4e6, 4e6))
        if rng.random() < 0.3:   # exact ties
            t = np.float32(np.floor(t) + 0.5)
        assert oracle.quantize_f32(float(t), 1.0) == closed_form(float(t))


def test_random_f64_against_exact(oracle):
    rng = np.random.default_rng(2)
    for _ in range(5000):
        x = float(rng.choice([-1, 1]) * 10 ** rng.uniform(-12, 12))
        r = 0.5 / float(10 ** rng.uniform(-8, 2))
        assert oracle.quantize_f64(x, r) == exact_q64(x, r)


def test_saturation_and_specials(oracle):
    cases = [np.inf, -np.inf, np.nan, 2.0**31, -(2.0**31), 3e9, -3e9, 2.0**31 - 128, -(2.0**31) + 128]
    for v in cases:
        assert oracle.quantize_f32(v, 1.0) == exact_q32(np.float32(v), np.float32(1.0)), v
        assert oracle.quantize_f64(v, 1.0) == exact_q64(v, 1.0), v
    assert oracle.quantize_f32(np.nan, 1.0) == -1                 # NaN: trunc 0, t >= -0.5 false
    assert oracle.quantize_f32(-np.inf, 1.0) == I32_MAX           # INT_MIN - 1 wraps
    assert oracle.quantize_f64(-2147483647.5, 1.0) == I32_MIN     # only f64 reaches INT_MIN
    assert oracle.quantize_f32(1e-45, 1.0) == 0                   # subnormal input


def test_fused_rounding_is_not_used(oracle):
    # For |t| < 2^22, RN(RN(x*r) + 0.5) == RN(x*r + 0.5) always; above 2^23 they
    # differ. x*r = 2^23 + 0.5 - 2^-24: unfused t = 2^23, t + 0.5 ties to even -> 2^23;
    # one fused rounding of x*r + 0.5 = 2^23 + 1 - 2^-24 gives 2^23 + 1.
    x, r = np.float32(8388609.0), np.float32(1 - 2.0**-24)
    assert oracle.quantize_f32(float(x), float(r)) == 8388608 == exact_q32(x, r)
    # f64: 0.5 stays on the grid up to 2^52 and codes saturate at 2^31, so fused
    # and unfused evaluation agree for every input; the saturated case:
    x, r = 2.0**53 + 2, 1 - 2.0**-53
    assert oracle.quantize_f64(x, r) == exact_q64(x, r) == I32_MAX


def test_reciprocal_rounding(oracle):
    for eb in [1e-3, 3e-7, 0.1, 7.0]:
        assert oracle.lib().cuszp_oracle_recip_f32(eb) == rn32(0.5 / float(np.float32(eb)))
        assert oracle.lib().cuszp_oracle_recip_f64(eb) == 0.5 / eb
