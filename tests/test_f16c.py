"""FP16C storage codec (1 sign, 4 exponent, 11 mantissa bits; the device's
StoreF16C, restated in the oracle) pinned by the definition of the format:
value = (-1)^s 2^(e-15) (1 + m/2^11) for e >= 1, (-1)^s 2^-14 (m/2^11) for
e = 0; encoding rounds to the nearest code (ties to even mantissa),
saturates at the largest code and sends NaN to the largest negative code."""
import math

import numpy as np

from oracle import oracle as orc


def value_of(code):
    s = -1.0 if code & 0x8000 else 1.0
    e = (code >> 11) & 0xF
    m = code & 0x7FF
    if e == 0:
        return s * math.ldexp(m / 2048.0, -14)
    return s * math.ldexp(1.0 + m / 2048.0, e - 15)


def test_decode_matches_format_definition_for_every_code():
    for code in range(0, 1 << 16, 1):
        assert orc.f16c_decode(code) == value_of(code), hex(code)


def test_encode_decode_round_trip_every_code():
    for code in range(0, 1 << 16):
        v = value_of(code)
        got = orc.f16c_encode(v)
        if v == 0.0:
            assert got in (0, 0x8000)
        else:
            assert got == code, (hex(code), hex(got))


def test_encode_is_nearest_code_ties_to_even():
    rng = np.random.default_rng(7)
    mags = 10.0 ** rng.uniform(-9, math.log10(1.99), 20000)
    vals = np.concatenate([mags, -mags]).astype(np.float32)
    for v in vals:
        c = orc.f16c_encode(float(v))
        d = orc.f16c_decode(c)
        mag = c & 0x7FFF
        for nb in (mag - 1, mag + 1):
            if 0 <= nb <= 0x7FFF:
                alt = value_of((c & 0x8000) | nb)
                assert abs(d - float(v)) <= abs(alt - float(v)), (float(v), hex(c))
    # exact midpoints between neighbouring codes go to the even mantissa
    for code in (0x0801, 0x3C00, 0x4002, 0x7001):
        lo, hi = value_of(code), value_of(code + 1)
        mid = np.float32((lo + hi) / 2.0)
        if float(mid) == (lo + hi) / 2.0:
            want = code if code % 2 == 0 else code + 1
            assert orc.f16c_encode(float(mid)) == want


def test_saturation_and_relative_precision():
    assert orc.f16c_encode(5.0) == 0x7FFF
    assert orc.f16c_encode(-5.0) == 0xFFFF
    assert orc.f16c_encode(float("inf")) == 0x7FFF
    assert orc.f16c_encode(float("-inf")) == 0xFFFF
    assert orc.f16c_encode(float("nan")) == 0xFFFF
    assert orc.f16c_encode(value_of(0x7FFF)) == 0x7FFF
    assert orc.f16c_encode(-value_of(0x7FFF)) == 0xFFFF
    # relative rounding error of normal values is at most 2^-12
    rng = np.random.default_rng(3)
    for v in rng.uniform(2.0 ** -14, 1.99, 5000).astype(np.float32):
        d = orc.f16c_decode(orc.f16c_encode(float(v)))
        assert abs(d - float(v)) <= float(v) * 2.0 ** -12 * (1 + 1e-6)
