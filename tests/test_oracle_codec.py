"""Pins the oracle's fast_prng and compressed_float3 codec (tensors.h:276-394).

Known answers are the values the survey recorded from the reference headers
compiled for the host; the rest are closed forms and invariants of the
format (5-bit shared exponent, offset 15, sign + 8-bit mantissa lanes).
"""
import math

import numpy as np
import pytest

SEED = 12894712  # fast_prng default (tensors.h:280)


def test_prng_known_answers(orc):
    assert orc.prng_ints(SEED, 4) == [-519354323, -1038438565, 2137405779, 762840351]
    f = orc.prng_floats(SEED, 3)
    # the survey printed these three in reverse call order
    assert np.allclose(f, [-0.120921604, -0.241780326, 0.497653574], rtol=0, atol=5e-9)
    # generate_float is generate_int * 2.3283064e-10 in float32
    ints = orc.prng_ints(SEED, 3)
    assert [np.float32(np.float32(i) * np.float32(2.3283064e-10)) for i in ints] == [np.float32(v) for v in f]


def test_prng_closed_form(orc):
    """One step: s += 1124781675; s ^= s >> 7 (arithmetic); s *= 1246477263, mod 2^32."""
    def step(s):
        s = (s + 1124781675) & 0xFFFFFFFF
        s = s - (1 << 32) if s >= 1 << 31 else s
        s ^= s >> 7
        s = (s * 1246477263) & 0xFFFFFFFF
        return s - (1 << 32) if s >= 1 << 31 else s
    for seed in (0, 1, -1, 12345, SEED, 2 ** 31 - 1, -2 ** 31):
        s, out = seed, []
        for _ in range(5):
            s = step(s)
            out.append(s)
        assert orc.prng_ints(seed, 5) == out


def test_thread_seed_known_answer(orc):
    assert orc.thread_seed(777, 3, 5, 70) == 1448313939


def test_cf3_known_answers(orc):
    assert orc.cf3_pack(1.0, -0.5, 0.25) == 0x78828080
    assert orc.cf3_pack(math.pi, math.e, -math.sqrt(2)) == 0x856D5CC9
    # e = 31 step is 512: 1e5 -> 99840; the small lanes flush to +0 and -0
    assert orc.cf3_pack(1e5, 1.0, -1.0) == 0xFC0000C3
    assert orc.cf3_unpack(0xFC0000C3)[0].tolist() == [99840.0, 0.0, -0.0]
    # sign bits with zero mantissa: -0 is preserved
    w, _ = orc.cf3_pack_dither(0.0, 0.0, 0.0, 12345)
    assert w == 0x04020000
    assert orc.cf3_unpack(0xFFFFFFFF)[0].tolist() == [-130560.0] * 3
    assert orc.cf3_unpack(1)[0].tolist() == [2.0 ** -22, 0.0, 0.0]


def test_cf3_unpack_closed_form(orc):
    rng = np.random.default_rng(1)
    w = rng.integers(0, 2 ** 32, size=20000, dtype=np.uint64).astype(np.uint32)
    got = orc.cf3_unpack(w).astype(np.float64)
    e = (w >> 27).astype(np.int64)
    for lane, (mpos, spos) in enumerate(((0, 8), (9, 17), (18, 26))):
        m = ((w >> mpos) & 255).astype(np.float64)
        s = ((w >> spos) & 1).astype(np.float64)
        want = m * (1 - 2 * s) * np.exp2(e - 22)
        assert np.array_equal(got[:, lane], want)
        # signed zero lanes
        zero = m == 0
        assert np.array_equal(np.signbit(got[zero, lane]), s[zero] == 1)


def test_cf3_roundtrip_on_grid(orc):
    """Values on the 8-bit grid of their shared exponent survive a round trip,
    dithered or not (the dither amplitude 0.95/2 < 1/2)."""
    rng = np.random.default_rng(2)
    for _ in range(300):
        e = int(rng.integers(1, 31))
        m = rng.integers(0, 256, size=3)
        m[int(rng.integers(0, 3))] = int(rng.integers(128, 256))  # the lane that sets the exponent
        sgn = rng.choice([-1.0, 1.0], size=3)
        v = (m * sgn * 2.0 ** (e - 22)).astype(np.float32)
        w = orc.cf3_pack(*v)
        assert np.array_equal(orc.cf3_unpack(w)[0], v)
        wd, _ = orc.cf3_pack_dither(*v, int(rng.integers(-2 ** 31, 2 ** 31)))
        assert np.array_equal(orc.cf3_unpack(wd)[0], v)
        assert (w >> 27) == e


def test_cf3_exponent_rule(orc):
    """e = clamp(max biased exponent - 112, 0, 31), then r = v * 2^(22-e)."""
    rng = np.random.default_rng(3)
    v = (rng.standard_normal((5000, 3)) * np.exp2(rng.integers(-30, 20, size=(5000, 1)))).astype(np.float32)
    w = orc.cf3_pack_array(v)
    bexp = ((v.view(np.uint32) >> 23) & 255).max(axis=1).astype(np.int64)
    e = np.clip(bexp - 112, 0, 31)
    assert np.array_equal(w >> 27, e.astype(np.uint32))
    r = v.astype(np.float64) * np.exp2(22 - e)[:, None]
    m = np.floor(np.minimum(np.abs(r), 255.0) + 0.5)
    dec = orc.cf3_unpack(w).astype(np.float64)
    assert np.array_equal(np.abs(dec) * np.exp2(22 - e)[:, None], m)


def test_cf3_dither_unbiased(orc):
    """The dithered pack is unbiased: the mean of many packs of a value
    between two grid points approaches the value; the plain pack does not."""
    x = np.float32(1.0 + 0.3 / 128)  # 0.3 of a grid step above 1.0 (e such that step is 1/128)
    dec = []
    for seed in range(4000):
        w, _ = orc.cf3_pack_dither(x, 0.5, 0.25, seed * 7919 + 1)
        dec.append(orc.cf3_unpack(w)[0, 0])
    mean = float(np.mean(dec))
    assert abs(mean - float(x)) < 0.05 / 128
    assert orc.cf3_unpack(orc.cf3_pack(x, 0.5, 0.25))[0, 0] == 1.0


def test_cf3_dither_draw_order(orc):
    """Three draws per pack, x then y then z (tensors.h:365-367)."""
    seed = 99
    _, s_after = orc.cf3_pack_dither(0.1, 0.2, 0.3, seed)
    assert s_after == orc.prng_ints(seed, 3)[-1]


@pytest.mark.parametrize("val", [np.nan, np.inf, -np.inf])
def test_cf3_non_finite(orc, val):
    """NaN and +-inf saturate through the clamp (fminf/fmaxf drop NaN)."""
    w = orc.cf3_pack(val, 1.0, 1.0)
    assert w >> 27 == 31
    lane = int(w & 255), int((w >> 8) & 1)
    assert lane == ((255, 1) if val == -np.inf else (255, 0))
