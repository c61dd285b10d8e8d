"""Philox4x32-10 and the FP32 primitives of the engine convention.

Known-answer vectors are the published ones of Salmon et al. (SC'11) as
distributed with Random123 (kat_vectors: philox4x32 10 rounds)."""
import math

import numpy as np
import pytest

from oracle import mcref

KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,expect", KAT)
def test_philox_kat(ctr, key, expect):
    assert tuple(int(v) for v in mcref.philox(ctr, key)) == expect


def test_exp_neg_accuracy():
    xs = np.concatenate([-np.logspace(-8, math.log10(79.0), 4000), [0.0, -1e-30, -0.5, -math.log(2)]])
    for x in xs:
        x32 = float(np.float32(x))
        got = mcref.exp_neg(x32)
        want = math.exp(x32)
        assert abs(got - want) <= 2.5e-7 * want + 1e-38, (x32, got, want)
    assert mcref.exp_neg(-81.0) == 0.0


def test_sincos_unit_circle_and_octants():
    rng = np.random.default_rng(0)
    for v in list(rng.integers(0, 2 ** 32, 3000)) + [0, 2 ** 29, 2 ** 30, 2 ** 31, 3 * 2 ** 29, 2 ** 32 - 1]:
        c, s = mcref.sincos(int(v))
        phi = 2 * math.pi * (int(v) >> 5) / 2 ** 27
        assert abs(c - math.cos(phi)) < 3e-7 and abs(s - math.sin(phi)) < 3e-7, (v, c, s, phi)


def test_f16_round_matches_numpy():
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.normal(size=5000).astype(np.float32), np.float32([1e-6, -3e-5, 65504, 0.1, 1 / 3])])
    for x in xs:
        assert mcref.f16_round(float(x)) == float(np.float32(np.float16(x))), x
