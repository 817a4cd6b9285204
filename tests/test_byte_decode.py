"""The kernels decode a unorm8 byte (texture.rs:173-175: u8 / 255.0) without
a division or a table: q = b * RN(1/255), then one fused correction
q + fma(-q, 255, b) * RN(1/255) (wgc_core.h `u8f`).  This checks, in exact
rational arithmetic with binary32 rounding of each fused step, that the
result equals the correctly rounded quotient for every byte."""
from fractions import Fraction

import numpy as np


def round_f32(x: Fraction) -> Fraction:
    """Round a rational to the nearest binary32 (ties to even)."""
    if x == 0:
        return Fraction(0)
    sign = -1 if x < 0 else 1
    x = abs(x)
    e = 0
    while x >= 2:
        x /= 2
        e += 1
    while x < 1:
        x *= 2
        e -= 1
    m = x * (1 << 23)
    fl = m.numerator // m.denominator
    rem = m - fl
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and fl % 2 == 1):
        fl += 1
    return sign * Fraction(fl, 1 << 23) * Fraction(2) ** e


def test_fused_reciprocal_decode_is_the_division_for_every_byte():
    r = round_f32(Fraction(1, 255))
    assert float(r) == float(np.float32(1.0) / np.float32(255.0))
    for b in range(256):
        want = round_f32(Fraction(b, 255))
        q = round_f32(b * r)
        e = round_f32(Fraction(b) - q * 255)
        got = round_f32(e * r + q)
        assert got == want, b
        assert float(got) == float(np.float32(b) / np.float32(255.0))
