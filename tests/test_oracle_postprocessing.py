"""Pinning the post-processing oracle (oracle/mqpiv_oracle_pp.c) against closed
forms and independent library routines (numpy's median/mean/std), on the CPU.
Reference: src/postprocessing/postprocessing.jl."""
from __future__ import annotations

import math
from fractions import Fraction

import numpy as np
import pytest

from multi_quickpiv_amd.postprocessing import cosd
from oracle import oracle as O


def _rand_field(shape, seed, dtype=np.float64, zeros=0):
    rng = np.random.default_rng(seed)
    VF = rng.normal(size=shape).astype(dtype)
    flat = VF.reshape(shape[0], -1)
    for i in rng.choice(flat.shape[1], size=zeros, replace=False):
        flat[:, i] = 0
    return np.asfortranarray(VF)


def _box(shape, idx, r):
    return tuple(slice(max(0, i - rr), min(s, i + rr + 1)) for i, rr, s in zip(idx, r, shape))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("shape", [(2, 9, 7), (3, 5, 4, 6)])
def test_magnitudes_match_numpy(dtype, shape):
    VF = _rand_field(shape, 1, dtype)
    M = O.pp_magnitudes(VF)
    assert M.dtype == dtype
    assert np.array_equal(M, np.sqrt(np.sum(VF * VF, axis=0, dtype=dtype)))


def test_th_and_std_masks_match_numpy():
    VF = _rand_field((2, 30, 20), 2)
    M = np.sqrt((VF ** 2).sum(0))
    assert np.array_equal(O.pp_th_mask(M, 1.5, 0), (M > 1.5).astype(np.uint8))
    assert np.array_equal(O.pp_th_mask(M, 1.5, 1), (M < 1.5).astype(np.uint8))
    mask, mean, sd = O.pp_std_mask(M, 1.2)
    assert mean == pytest.approx(M.mean(), rel=1e-14)
    assert sd == pytest.approx(M.std(ddof=1), rel=1e-13)  # corrected std (Statistics.std)
    assert np.array_equal(mask, (np.abs(M - M.mean()) > 1.2 * M.std(ddof=1)).astype(np.uint8))
    # one value: corrected std is NaN, nothing is flagged
    m1, _, sd1 = O.pp_std_mask(np.array([3.0]), 1.0)
    assert math.isnan(sd1) and m1.tolist() == [0]


@pytest.mark.parametrize("shape,r", [((2, 9, 7), (2, 2)), ((2, 6, 5), (1, 3)), ((3, 5, 4, 6), (1, 2, 1))])
def test_median_replacement_matches_numpy(shape, r):
    VF = _rand_field(shape, 3)
    rng = np.random.default_rng(4)
    mask = (rng.random(shape[1:]) < 0.3).astype(np.uint8)
    out = O.pp_replace(VF, mask, 1, r)
    zero = O.pp_replace(VF, mask, 0, r)
    for idx in np.ndindex(*shape[1:]):
        if not mask[idx]:
            assert np.array_equal(out[(slice(None),) + idx], VF[(slice(None),) + idx])
            continue
        assert np.all(zero[(slice(None),) + idx] == 0)
        b = _box(shape[1:], idx, r)
        for c in range(shape[0]):
            assert out[(c,) + idx] == np.median(VF[(c,) + b])


def test_median_even_count_nan_and_signed_zero():
    # 2x2 corner box (radius 1 at (0,0) of a 2x2 field): median of 4 = mean of the middle pair
    VF = np.asfortranarray(np.array([[[1.0, 10.0], [4.0, 2.0]], [[0.0, 0.0], [0.0, 0.0]]]))
    out = O.pp_replace(VF, np.ones((2, 2), np.uint8), 1, (1, 1))
    assert out[0, 0, 0] == 3.0  # sorted 1, 2, 4, 10 -> (2 + 4) / 2
    VF[1, 1, 1] = np.nan
    out = O.pp_replace(VF, np.ones((2, 2), np.uint8), 1, (1, 1))
    assert np.isnan(out[1, 0, 0]) and out[0, 0, 0] == 3.0
    # -0.0 sorts before +0.0 (Base.isless): median of (-0, -0, +0) is -0.0
    VF3 = np.asfortranarray(np.zeros((2, 3, 1)))
    VF3[0, 0, 0] = VF3[0, 1, 0] = -0.0
    out = O.pp_replace(VF3, np.array([[0], [1], [0]], np.uint8), 1, (1, 0))
    assert out[0, 1, 0] == 0 and math.copysign(1, out[0, 1, 0]) == -1


def test_average_matches_numpy_mean():
    VF = _rand_field((2, 11, 8), 5)
    out = O.pp_average(VF, (2, 1))
    for idx in np.ndindex(11, 8):
        b = _box((11, 8), idx, (2, 1))
        for c in range(2):
            assert out[(c,) + idx] == pytest.approx(np.mean(VF[(c,) + b]), rel=1e-14, abs=1e-15)
    const = np.asfortranarray(np.full((3, 4, 5, 3), 0.75))
    assert np.array_equal(O.pp_average(const, 2), const)


def _checker(n0, n1, a=(1.0, 0.0), b=(0.0, 2.0)):
    VF = np.zeros((2, n0, n1))
    for i in range(n0):
        for j in range(n1):
            VF[:, i, j] = a if (i + j) % 2 == 0 else b
    return np.asfortranarray(VF)


def test_collectiveness_closed_forms():
    md = cosd(40)
    uni = np.asfortranarray(np.tile(np.array([0.3, -1.1])[:, None, None], (1, 6, 5)))
    assert np.all(O.pp_collectiveness(uni, 2, md) == 1.0)
    # checkerboard of orthogonal directions: an interior 3x3 box holds 5 cells of
    # the centre's colour (itself + 4 diagonals)
    col = O.pp_collectiveness(_checker(7, 7), 1, md)
    assert np.all(col[1:-1, 1:-1] == np.float32(5 / 9))
    assert col[0, 0] == np.float32(2 / 4)  # corner: itself + 1 diagonal of 4
    # zero vectors normalise to NaN: similar to nothing, counted by nobody
    z = uni.copy(order="F")
    z[:, 2, 2] = 0
    cz = O.pp_collectiveness(z, 1, md)
    assert cz[2, 2] == 0 and cz[1, 1] == np.float32(8 / 9) and cz[0, 0] == 1.0
    # the angle test is strict and on the normalised dot product
    ang = np.deg2rad([0.0, 30.0])
    two = np.asfortranarray(np.stack([np.cos(ang), np.sin(ang)])[:, :, None])
    assert np.all(O.pp_collectiveness(two, 1, cosd(40)) == 1.0)
    assert np.all(O.pp_collectiveness(two, 1, cosd(20)) == 0.5)


def test_similarity_averaging_closed_forms():
    md = cosd(20)
    v = np.array([0.6, -0.8]) * 3.0
    uni = np.asfortranarray(np.tile(v[:, None, None], (1, 5, 4)))
    out = O.pp_simavg(uni, 1, md)
    assert np.allclose(out, (v / 3.0)[:, None, None], rtol=0, atol=1e-15)  # unit vectors, (9/9)^2
    chk = _checker(7, 7)
    out = O.pp_simavg(chk, 1, md)  # interior a-cell: 5 similar (1,0) -> (1,0) * (5/9)^2
    assert out[0, 3, 3] == pytest.approx((5 / 9) ** 2, rel=1e-15) and out[1, 3, 3] == 0
    assert out[1, 3, 4] == pytest.approx((5 / 9) ** 2, rel=1e-15)  # b-cell, (0,2) normalised
    raw = O.pp_simavg(chk, 1, md, normalize=False, scale=False)
    assert raw[0, 3, 3] == 5.0 and raw[1, 3, 4] == 10.0  # sum of similar vectors
    e3 = O.pp_simavg(chk, 1, md, normalize=True, scale=True, expo=3)
    assert e3[0, 3, 3] == pytest.approx((5 / 9) ** 3, rel=1e-15)
    z = uni.copy(order="F")
    z[:, 1, 1] = 0
    assert np.array_equal(O.pp_simavg(z, 1, md)[:, 1, 1], [0.0, 0.0])  # zero vectors are copied


def test_local_magnitude_mask_closed_form():
    VF = np.asfortranarray(np.tile(np.array([1.0, 0.0])[:, None, None], (1, 5, 5)))
    VF[:, 2, 2] = [-6.0, 8.0]  # magnitude 10 among ones
    out, removed = O.pp_lmm(VF, 3.0, 1)
    # the spike: 10 / mean(8 ones) = 10 >= 3 -> removed, set to signed zero
    assert removed[2, 2] == 1 and removed.sum() == 1
    assert out[0, 2, 2] == 0 and math.copysign(1, out[0, 2, 2]) == -1 and out[1, 2, 2] == 0
    # a neighbour: 1 / ((7 + 10) / 8) < 3 -> kept unchanged
    assert np.array_equal(out[:, 1, 1], [1.0, 0.0])


def test_similarity_exponent_power_is_faithful():
    """(ns/nn)^e through a field whose similar sum is exactly (1, 0): the centre
    of [(0,1), (1,0), (0,1)] has ns/nn = 1/3, so the output is the power itself,
    within one ulp of the exact power of the rounded base (exact for e <= 2)."""
    VF = np.asfortranarray(np.array([[0.0, 1.0, 0.0], [1.0, 0.0, 1.0]])[:, :, None])
    x = 1 / 3
    for e in range(0, 10):
        got = O.pp_simavg(VF, (1, 0), cosd(20), normalize=False, scale=True, expo=e)[0, 1, 0]
        ref = Fraction(x) ** e
        tol = Fraction(0) if e <= 2 else Fraction(math.ulp(float(ref)))
        assert abs(Fraction(got) - ref) <= tol + Fraction(math.ulp(float(ref))) / 2, e
