"""Philox4x32-10 draws: an independent Python statement of the published round
function (Salmon et al., SC'11) checked against the Random123 known-answer
vectors, and the oracle's C implementation and draw mapping against it."""
import math

import numpy as np
import pytest

M0, M1 = 0xD2511F53, 0xCD9E8D57
W0, W1 = 0x9E3779B9, 0xBB67AE85
MASK = 0xFFFFFFFF

# Random123 kat_vectors, philox4x32 with 10 rounds: (counter, key, output)
KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((MASK,) * 4, (MASK, MASK), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


def philox_py(ctr, key):
    c0, c1, c2, c3 = ctr
    k0, k1 = key
    for _ in range(10):
        p0, p1 = M0 * c0, M1 * c2
        c0, c1, c2, c3 = ((p1 >> 32) ^ c1 ^ k0) & MASK, p1 & MASK, ((p0 >> 32) ^ c3 ^ k1) & MASK, p0 & MASK
        k0, k1 = (k0 + W0) & MASK, (k1 + W1) & MASK
    return (c0, c1, c2, c3)


@pytest.mark.parametrize("ctr,key,out", KAT)
def test_python_philox_kat(ctr, key, out):
    assert philox_py(ctr, key) == out


@pytest.mark.parametrize("ctr,key,out", KAT)
def test_oracle_philox_kat(orc, ctr, key, out):
    assert tuple(orc.philox(ctr, key)) == out


def test_oracle_philox_random_counters(orc):
    rng = np.random.default_rng(1)
    for _ in range(200):
        ctr = [int(x) for x in rng.integers(0, 2**32, 4)]
        key = [int(x) for x in rng.integers(0, 2**32, 2)]
        assert tuple(orc.philox(ctr, key)) == philox_py(ctr, key)


def draw_py(seed, vox, jump, slot):
    x = philox_py((vox & MASK, vox >> 32, jump, slot), (seed & MASK, seed >> 32))
    s24 = 1.0 / 16777216.0
    u1 = ((x[0] >> 8) + 1) * s24
    u2 = (x[1] >> 8) * s24
    u3 = ((x[2] >> 8) + 1) * s24
    return math.sqrt(-2.0 * math.log(u1)) * math.cos(2 * math.pi * u2), math.log(u3)


def test_draw_mapping(orc):
    for seed, vox, jump, slot in [(8665904, 0, 0, 0), (8665904, 123456789012, 2249, 14), (2**40 + 7, 5, 17, 3)]:
        n, lu = orc.draw(seed, vox, jump, slot)
        n2, lu2 = draw_py(seed, vox, jump, slot)
        assert n == pytest.approx(n2, rel=1e-12, abs=1e-12)
        assert lu == pytest.approx(lu2, rel=1e-12)


def test_draw_moments(orc):
    ns = np.array([orc.draw(8665904, v, 0, 2)[0] for v in range(4000)])
    us = np.exp([orc.draw(8665904, v, 0, 2)[1] for v in range(4000)])
    assert abs(ns.mean()) < 0.06 and abs(ns.std() - 1.0) < 0.04
    assert abs(us.mean() - 0.5) < 0.02 and us.min() > 0 and us.max() <= 1.0
