"""Pins of the oracle's Owen-scrambled Sobol' sampler (shaders/RNG.hlsli:9-99):
closed forms of both Sobol' axes, the (0, k, 2)-net property of the unscrambled
points, of the Owen-scrambled points and of the sampler's stream over sample
indices, and the nesting property that makes the Laine-Karras hash an Owen
scramble (an input bit never changes a lower output bit).  The hash constants
themselves are those of the reference's listing; no known-answer vector for
them exists in the reference."""
import numpy as np
import pytest

import orc

SOBOL2, LK, OWEN, HASH = 1, 2, 3, 0


def _is_net(x, y, k):
    # every elementary box of area 2^-k holds exactly one of the 2^k points
    x = np.asarray(x, np.uint64)
    y = np.asarray(y, np.uint64)
    for m in range(k + 1):
        bx = x >> np.uint64(32 - m) if m else np.zeros_like(x)
        by = y >> np.uint64(32 - (k - m)) if k - m else np.zeros_like(y)
        cells = bx * np.uint64(1 << (k - m)) + by
        if len(np.unique(cells)) != 1 << k:
            return False
    return True


def test_axis0_is_bit_reversal():
    i = np.arange(4096, dtype=np.uint32)
    rev = np.array([int(f"{v:032b}"[::-1], 2) for v in i], np.uint32)
    assert np.array_equal(orc.sobol_part(SOBOL2, i, 0), rev)


def test_axis1_from_direction_numbers():
    # Sobol's second dimension: direction numbers v_k = row k-1 of Pascal's
    # triangle mod 2 read as a binary fraction (1/2, 3/4, 5/8, 15/16, 17/32,
    # ...); point i is the xor of v_k over the set bits k of i (binary index
    # order, as the reference's loop takes it, not Gray-code order)
    from math import comb
    v = [sum((comb(k, j) % 2) << (31 - j) for j in range(k + 1)) for k in range(32)]
    i = np.arange(1 << 12, dtype=np.uint32)
    expect = np.zeros_like(i)
    for k in range(12):
        expect ^= np.where((i >> k) & 1, np.uint32(v[k]), np.uint32(0))
    assert np.array_equal(orc.sobol_part(SOBOL2, i, 1), expect)
    first = expect[:8].astype(np.float64) / 2 ** 32
    assert np.array_equal(first, [0, 0.5, 0.75, 0.25, 0.625, 0.125, 0.375, 0.875])


@pytest.mark.parametrize("k", [1, 4, 7, 10])
def test_sobol_points_form_nets(k):
    i = np.arange(1 << k, dtype=np.uint32)
    assert _is_net(orc.sobol_part(SOBOL2, i, 0), orc.sobol_part(SOBOL2, i, 1), k)


def test_a_broken_axis_is_caught():
    # the net check is not vacuous: axis 1 with the van der Corput numbers
    # shifted by one bit is not a net
    i = np.arange(256, dtype=np.uint32)
    assert not _is_net(orc.sobol_part(SOBOL2, i, 0), orc.sobol_part(SOBOL2, i >> 1, 0), 8)


@pytest.mark.parametrize("seeds", [(1, 2), (0x9e3779b9, 0x85ebca6b), (123456789, 987654321)])
def test_owen_scrambling_keeps_the_net(seeds):
    k = 9
    i = np.arange(1 << k, dtype=np.uint32)
    x = orc.sobol_part(OWEN, orc.sobol_part(SOBOL2, i, 0), seeds[0])
    y = orc.sobol_part(OWEN, orc.sobol_part(SOBOL2, i, 1), seeds[1])
    assert _is_net(x, y, k)
    assert not np.array_equal(x, orc.sobol_part(SOBOL2, i, 0))


def test_laine_karras_is_nested():
    # flipping input bit j never changes output bits below j
    rng = np.random.default_rng(4)
    x = rng.integers(0, 2 ** 32, 4000, dtype=np.uint64).astype(np.uint32)
    seed = rng.integers(0, 2 ** 32, 4000, dtype=np.uint64).astype(np.uint32)
    base = orc.sobol_part(LK, x, seed)
    for j in (0, 5, 13, 22, 31):
        flipped = orc.sobol_part(LK, x ^ np.uint32(1 << j), seed)
        diff = base ^ flipped
        assert np.all(diff & np.uint32((1 << j) - 1) == 0)
        assert np.all(diff & np.uint32(1 << j))


@pytest.mark.parametrize("pixel,pad", [(0, 0), (777, 3), (123456, 11)])
def test_sampler_stream_forms_nets_over_samples(pixel, pad):
    # dimensions (2 pad, 2 pad + 1) of samples 0 .. 2^k - 1 of one pixel
    k = 8
    pts = np.array([orc.sobol(pixel, smp, 2 * pad, 2) for smp in range(1 << k)])
    assert _is_net(pts[:, 0], pts[:, 1], k)


def test_sampler_floats_in_unit_interval():
    v = np.concatenate([orc.sobol(p, s, 0, 64) for p in range(8) for s in range(8)]).astype(np.float64)
    f = np.minimum(np.float32(v) * np.float32(2.3283064365386963e-10), np.float32(0.99999994))
    assert f.min() >= 0 and f.max() < 1
    # 4096 draws of many dimensions: mean 1/2 within 5 sigma
    assert abs(f.mean() - 0.5) < 5 * np.sqrt(1 / 12 / len(f))
