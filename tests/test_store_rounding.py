"""Pins of the oracle's restatement of the 16-bit store codecs
(orc.store_codes / orc.decode_codes): round to nearest even against the C
oracle's FP16C encoder, numpy's float16 conversion and hand-worked BF16
ties; stochastic rounding against what its definition fixes exactly (over
every value of the random bits a codec uses, the mean decoded value is the
input), with a planted mistake in the restatement's inputs to show the
check has teeth. The device codec is compared with this restatement bit for
bit in tests/test_gpu_parity.py."""
import numpy as np
import pytest

from oracle import oracle as orc

RNG = np.random.default_rng(2024)
# random bits a codec uses: the top 12 (FP16C), 13 (FP16), 16 (BF16)
USED = {"f16c": 12, "f16": 13, "bf16": 16}


def sample(storage, n):
    """Finite values inside each format's normal range (FP16 holds the
    2^15-scaled deviation, so its inputs are larger)."""
    lo, hi = {"f16c": (-13, 0), "f16": (-13, 15), "bf16": (-30, 30)}[storage]
    mag = np.exp2(RNG.uniform(lo, hi, n))
    return (np.where(RNG.random(n) < 0.5, -1.0, 1.0) * mag).astype(np.float32)


def test_round_to_nearest_matches_the_c_encoder_and_the_format_rules():
    v = np.concatenate([sample("f16c", 3000), np.float32([0.0, -0.0, 2.0 ** -20, -(2.0 ** -25), 3.0, -7.0,
                                                          np.inf, -np.inf])]).astype(np.float32)
    zeros = np.zeros(v.size, np.uint32)
    got = orc.store_codes("f16c", v, zeros, False)
    want = np.array([orc.f16c_encode(float(x)) for x in v], dtype=np.uint16)
    assert np.array_equal(got, want)
    h = sample("f16", 3000)
    assert np.array_equal(orc.store_codes("f16", h, np.zeros(h.size, np.uint32), False),
                          h.astype(np.float16).view(np.uint16))
    # BF16 ties to even: 1 + 2^-8 -> 1.0, 1 + 3 * 2^-8 -> 1 + 2^-6; just above a tie rounds up
    b = np.float32([1 + 2.0 ** -8, 1 + 3 * 2.0 ** -8, 1 + 2.0 ** -8 + 2.0 ** -20, -(1 + 2.0 ** -8)])
    assert orc.store_codes("bf16", b, np.zeros(4, np.uint32), False).tolist() == [0x3F80, 0x3F82, 0x3F81, 0xBF80]


@pytest.mark.parametrize("storage", ["f16c", "f16", "bf16"])
def test_stochastic_rounding_mean_is_exact(storage):
    """Over all 2^k values of the k random bits used, the decoded values
    average to the input exactly, and every result is one of the two codes
    that bracket it."""
    k = USED[storage]
    n = 40 if storage == "bf16" else 200
    v = sample(storage, n)
    r = (np.arange(2 ** k, dtype=np.uint32) << np.uint32(32 - k))
    codes = orc.store_codes(storage, np.repeat(v, r.size), np.tile(r, n), True).reshape(n, -1)
    dec = orc.decode_codes(storage, codes.reshape(-1)).reshape(n, -1)
    assert np.array_equal(dec.mean(axis=1), v.astype(np.float64))
    nearest = orc.decode_codes(storage, orc.store_codes(storage, v, np.zeros(n, np.uint32), False))
    lo = np.minimum(dec.min(axis=1), nearest)
    hi = np.maximum(dec.max(axis=1), nearest)
    assert np.all(lo <= v) and np.all(v <= hi)
    assert all(np.unique(c).size <= 2 for c in codes)
    # representable inputs are stored exactly whatever the bits
    rep = orc.decode_codes(storage, orc.store_codes(storage, v, np.zeros(n, np.uint32), False)).astype(np.float32)
    again = orc.store_codes(storage, np.repeat(rep, 64), RNG.integers(0, 2 ** 32, 64 * n, dtype=np.uint64)
                            .astype(np.uint32), True)
    assert np.array_equal(orc.decode_codes(storage, again), np.repeat(rep, 64).astype(np.float64))


def test_stochastic_mean_check_has_teeth():
    """Feeding the FP16C codec random words with one bit too few (the
    shift of a 13-bit codec) biases the mean, which the exactness check
    catches."""
    v = sample("f16c", 50)
    r = (np.arange(4096, dtype=np.uint32) << np.uint32(20)) >> np.uint32(1)  # planted: half the range
    codes = orc.store_codes("f16c", np.repeat(v, r.size), np.tile(r, v.size), True)
    mean = orc.decode_codes("f16c", codes).reshape(v.size, -1).mean(axis=1)
    assert not np.array_equal(mean, v.astype(np.float64))
