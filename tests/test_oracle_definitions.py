"""Pins the oracle's masked NSQECC and ZNCC against their definitions, by brute
force in numpy on small windows, and shows that each check catches a plausible
misreading of the reference planted in the oracle (orc_set_plant).

Definitions (reference files relative to its repository root):

* masked NSQECC (mask_nsqecc.jl:141-160, masked_nsqecc_ops.jl:19-42, 123-143):
  with F the IA, m its mask window and G the zero-padded search window placed at
  SR_TLF_off, the reference forms C1 = corr(F*m, G), C2 = corr(m, G^2) and
  sum(F^2) = sum((F*m)^2 / w), w = m where m > 0 else 1.  For any mask m >= 0
  this is num(s) = sum_x m(x) (F(x) - G(x+s))^2 and
  den(s) = sqrt(sum_x m(x) G(x+s)^2) * sqrt(sum_x m(x) F(x)^2), S = den/(den+num).
  The brute force below evaluates that form directly; it shares no code with the
  oracle's transforms.
* ZNCC (zncc.jl:117-144, zncc_ops.jl:3-51): F' = F - mean(F), G' = G - mean of
  the clipped search region (padding stays 0), Z(s) = sum_x F'(x) G'(x+s) /
  (sqrt(sum F'^2) * sqrt(sum over the IA-sized window at s of G'^2)).

The geometry these windows use (coords) is pinned by closed forms in
test_oracle_pinning.py.
"""
from __future__ import annotations

import contextlib

import numpy as np
import pytest

from oracle import oracle as O

RTOL = 1e-10


@contextlib.contextmanager
def planted(k):
    O.set_plant(k)
    try:
        yield
    finally:
        O.set_plant(O.PLANT_NONE)


def _windows(a, b, m, p, lin):
    """(F, M, G, V): the IA of image 1 (and of the mask), and the search window
    of image 2 zero-padded to is + 2 SM with its clipped region placed at
    SR_TLF_off; V marks the positions that hold image data."""
    co = O.coords(p, a.shape, lin)
    tlf, sr_tlf, sr_brb, off = co[0] - 1, co[2] - 1, co[3] - 1, co[4]
    isz = (p.is_[0], p.is_[1])
    ss = tuple(isz[d] + 2 * p.sm[d] for d in range(2))
    ia = tuple(slice(tlf[d], tlf[d] + isz[d]) for d in range(2))
    F = a[ia]
    M = m[ia] if m is not None else None
    src = b[tuple(slice(sr_tlf[d], sr_brb[d] + 1) for d in range(2))]
    G = np.zeros(ss)
    V = np.zeros(ss, dtype=bool)
    dst = tuple(slice(off[d], off[d] + src.shape[d]) for d in range(2))
    G[dst] = src
    V[dst] = True
    return F, M, G, V


def _masked_def(F, M, G, sm):
    ns = (2 * sm[0] + 1, 2 * sm[1] + 1)
    sF = np.sum(M * F * F)
    S = np.empty(ns)
    for s0 in range(ns[0]):
        for s1 in range(ns[1]):
            Gw = G[s0:s0 + F.shape[0], s1:s1 + F.shape[1]]
            num = np.sum(M * (F - Gw) ** 2)
            den = np.sqrt(np.sum(M * Gw * Gw)) * np.sqrt(sF)
            S[s0, s1] = den / (den + num)
    return S


def _zncc_def(F, G, V, sm):
    ns = (2 * sm[0] + 1, 2 * sm[1] + 1)
    Fp = F - F.mean()
    Gp = np.where(V, G - G[V].mean(), 0.0)
    sF = np.sqrt(np.sum(Fp * Fp))
    Z = np.empty(ns)
    for s0 in range(ns[0]):
        for s1 in range(ns[1]):
            Gw = Gp[s0:s0 + F.shape[0], s1:s1 + F.shape[1]]
            Z[s0, s1] = np.sum(Fp * Gw) / (sF * np.sqrt(np.sum(Gw * Gw)))
    return Z


def _pair(shape, seed, d=(2, -3)):
    """A particle-like pair with a known shift plus independent noise, so the
    similarity has a clear match as well as unrelated positions."""
    r = np.random.default_rng(seed)
    a = 0.1 + 0.02 * r.standard_normal(shape)
    idx = tuple(r.integers(0, s, 40) for s in shape)
    a[idx] += r.uniform(0.5, 1.0, 40)
    b = np.roll(a, d, axis=(0, 1)) + 0.02 * r.standard_normal(shape)
    return a, b


# IAs of a 48 x 44 image at interSize 12, margin 5, step 8: the first (clipped,
# SR_TLF_off > 0), an interior one and the last (SR_BRB_off > 0)
_P = ("mask_nsqecc", 2, (12, 10), (5, 4), (8, 7))
_IAS = (1, 7, 12, 30)


def _masked_err(a, b, m, lin, mode):
    p = O.make_params(*_P)
    F, M, G, _ = _windows(a, b, m, p, lin)
    ref = _masked_def(F, M, G, (p.sm[0], p.sm[1]))
    got, _ = O.ia_similarity(a, b, p, lin, mask=m, mode=mode)
    return float(np.max(np.abs(got - ref) / np.abs(ref)))


def _binary_mask(shape, seed):
    return (np.random.default_rng(seed).random(shape) > 0.3).astype(np.float64)


def _fractional_mask(shape, seed):
    r = np.random.default_rng(seed)
    m = r.random(shape)
    m[r.random(shape) < 0.2] = 0.0
    return m


@pytest.mark.parametrize("mode", [O.DIRECT, O.FFT])
@pytest.mark.parametrize("kind", ["binary", "fractional"])
def test_masked_nsqecc_matches_definition(oracle, mode, kind):
    a, b = _pair((48, 44), 11)
    m = _binary_mask(a.shape, 12) if kind == "binary" else _fractional_mask(a.shape, 13)
    for lin in _IAS:
        assert _masked_err(a, b, m, lin, mode) < RTOL, lin


def _windows_nd(a, b, m, p, lin):
    """_windows for any ndims (3-D volumes)."""
    nd = a.ndim
    co = O.coords(p, a.shape, lin)
    tlf, sr_tlf, sr_brb, off = co[0] - 1, co[2] - 1, co[3] - 1, co[4]
    isz = tuple(p.is_[d] for d in range(nd))
    ss = tuple(isz[d] + 2 * p.sm[d] for d in range(nd))
    ia = tuple(slice(tlf[d], tlf[d] + isz[d]) for d in range(nd))
    src = b[tuple(slice(sr_tlf[d], sr_brb[d] + 1) for d in range(nd))]
    G = np.zeros(ss)
    G[tuple(slice(off[d], off[d] + src.shape[d]) for d in range(nd))] = src
    return a[ia], m[ia], G


def _masked_def_nd(F, M, G, sm):
    ns = tuple(2 * s + 1 for s in sm)
    sF = np.sum(M * F * F)
    S = np.empty(ns)
    for s in np.ndindex(*ns):
        Gw = G[tuple(slice(s[d], s[d] + F.shape[d]) for d in range(F.ndim))]
        num = np.sum(M * (F - Gw) ** 2)
        den = np.sqrt(np.sum(M * Gw * Gw)) * np.sqrt(sF)
        S[s] = den / (den + num)
    return S


# 3-D: a 20 x 18 x 16 volume at interSize (6, 5, 4), margin (2, 2, 1), step
# (5, 4, 4); IAs of the first (clipped) corner, the interior and the last corner
_P3 = ("mask_nsqecc", 3, (6, 5, 4), (2, 2, 1), (5, 4, 4))


@pytest.mark.parametrize("mode", [O.DIRECT, O.FFT])
@pytest.mark.parametrize("kind", ["binary", "fractional"])
def test_masked_nsqecc_3d_matches_definition(oracle, mode, kind):
    r = np.random.default_rng(21)
    shape = (20, 18, 16)
    a = 0.1 + 0.02 * r.standard_normal(shape)
    idx = tuple(r.integers(0, s, 60) for s in shape)
    a[idx] += r.uniform(0.5, 1.0, 60)
    b = np.roll(a, (1, -2, 1), axis=(0, 1, 2)) + 0.02 * r.standard_normal(shape)
    m = (r.random(shape) > 0.3).astype(np.float64)
    if kind == "fractional":
        m = r.random(shape) * (r.random(shape) > 0.2)
    p = O.make_params(*_P3)
    n = int(np.prod(O.vf_size(p, shape)))
    for lin in (1, n // 2, n):
        F, M, G = _windows_nd(a, b, m, p, lin)
        ref = _masked_def_nd(F, M, G, tuple(p.sm[d] for d in range(3)))
        got, _ = O.ia_similarity(a, b, p, lin, mask=m, mode=mode)
        assert got.shape == ref.shape
        assert float(np.max(np.abs(got - ref) / np.abs(ref))) < RTOL, lin


@pytest.mark.parametrize("plant", ["PLANT_MASK_C1_UNMASKED", "PLANT_MASK_C2_BOXSUM", "PLANT_MASK_NO_W"])
def test_masked_3d_definition_check_catches_planted_mistake(oracle, plant):
    r = np.random.default_rng(22)
    shape = (20, 18, 16)
    a = 0.1 + 0.02 * r.standard_normal(shape)
    b = np.roll(a, (1, -2, 1), axis=(0, 1, 2)) + 0.02 * r.standard_normal(shape)
    m = r.random(shape) * (r.random(shape) > 0.2)
    p = O.make_params(*_P3)
    lin = int(np.prod(O.vf_size(p, shape))) // 2
    F, M, G = _windows_nd(a, b, m, p, lin)
    ref = _masked_def_nd(F, M, G, tuple(p.sm[d] for d in range(3)))
    with planted(getattr(O, plant)):
        got, _ = O.ia_similarity(a, b, p, lin, mask=m, mode=O.DIRECT)
    assert float(np.max(np.abs(got - ref) / np.abs(ref))) > 1e-3


def test_masked_nsqecc_ignores_image1_where_mask_is_zero(oracle):
    a, b = _pair((48, 44), 14)
    m = _binary_mask(a.shape, 15)
    a2 = np.where(m == 0, a + np.random.default_rng(16).uniform(-5, 5, a.shape), a)
    p = O.make_params(*_P)
    for lin in _IAS:
        S1, _ = O.ia_similarity(a, b, p, lin, mask=m)
        S2, _ = O.ia_similarity(a2, b, p, lin, mask=m)
        np.testing.assert_allclose(S2, S1, rtol=1e-12, atol=0)


@pytest.mark.parametrize("plant,kind", [(O.PLANT_MASK_C1_UNMASKED, "binary"),
                                        (O.PLANT_MASK_C2_BOXSUM, "binary"),
                                        (O.PLANT_MASK_NO_W, "fractional")])
def test_masked_definition_check_catches_planted_mistake(oracle, plant, kind):
    a, b = _pair((48, 44), 11)
    m = _binary_mask(a.shape, 12) if kind == "binary" else _fractional_mask(a.shape, 13)
    with planted(plant):
        worst = max(_masked_err(a, b, m, lin, O.DIRECT) for lin in _IAS)
    assert worst > 1e-3, worst
    # and the unplanted oracle is back at the definition
    assert max(_masked_err(a, b, m, lin, O.DIRECT) for lin in _IAS) < RTOL


def test_masked_invariant_catches_unmasked_correlation(oracle):
    a, b = _pair((48, 44), 14)
    m = _binary_mask(a.shape, 15)
    a2 = np.where(m == 0, a + 3.0, a)
    p = O.make_params(*_P)
    with planted(O.PLANT_MASK_C1_UNMASKED):
        S1, _ = O.ia_similarity(a, b, p, 7, mask=m)
        S2, _ = O.ia_similarity(a2, b, p, 7, mask=m)
    assert np.max(np.abs(S2 - S1)) > 1e-3


_PZ = ("zncc", 2, (12, 10), (5, 4), (8, 7))


def _zncc_err(a, b, lin, mode):
    p = O.make_params(*_PZ)
    F, _, G, V = _windows(a, b, None, p, lin)
    ref = _zncc_def(F, G, V, (p.sm[0], p.sm[1]))
    got, _ = O.ia_similarity(a, b, p, lin, mode=mode)
    return float(np.max(np.abs(got - ref)))  # |Z| <= 1: absolute error


@pytest.mark.parametrize("mode", [O.DIRECT, O.FFT])
def test_zncc_matches_definition(oracle, mode):
    a, b = _pair((48, 44), 21)
    b = b + 0.7  # a search-region mean far from the IA's
    for lin in _IAS:
        assert _zncc_err(a, b, lin, mode) < RTOL, lin


def test_zncc_definition_check_catches_raw_sum_of_squares(oracle):
    a, b = _pair((48, 44), 21)
    b = b + 0.7
    with planted(O.PLANT_ZNCC_RAW_F2):
        worst = max(_zncc_err(a, b, lin, O.DIRECT) for lin in _IAS)
    assert worst > 1e-3, worst
    assert max(_zncc_err(a, b, lin, O.DIRECT) for lin in _IAS) < RTOL


def test_zncc_peak_is_one_for_an_exact_copy_with_offset(oracle):
    """Z = 1 where image 2 repeats the IA plus a constant, if the search region
    has the IA's mean offset by that constant: here image 2 is image 1 plus 0.5
    with the same mean-free search region (an integer shift of a periodic
    pattern), so the definition's two means differ by exactly 0.5."""
    t = np.arange(48)
    a = np.sin(2 * np.pi * t[:, None] / 12) * np.cos(2 * np.pi * t[None, :] / 8)
    b = np.roll(a, (3, 2), axis=(0, 1)) + 0.5
    p = O.make_params("zncc", 2, 12, 6, 12)  # interior IA 6: TLF 13, SR 7..30 (two periods)
    Z, co = O.ia_similarity(a, b, p, 6)
    assert co[4][0] == 0 and co[4][1] == 0
    assert Z[6 + 3, 6 + 2] == pytest.approx(1.0, abs=1e-12)


def _skip_def(img1, mask, p, masked):
    """skip_inter_region by brute force (piv_ops.jl:134-157, maxval filter)."""
    vf = O.vf_size(p, img1.shape)
    out = []
    for lin in range(int(np.prod(vf))):
        v = np.unravel_index(lin, vf, order="F")
        sl = tuple(slice(v[d] * p.st[d], v[d] * p.st[d] + p.is_[d]) for d in range(2))
        if masked:
            s = p.threshold > 0 and img1[sl].max() < p.threshold
            s = s or (p.mask_threshold > 0 and mask[sl].max() < p.mask_threshold)
        else:
            s = p.threshold >= 0 and img1[sl].max() < p.threshold
        out.append(int(s))
    return np.array(out, dtype=np.uint8)


@pytest.mark.parametrize("masked,thr,mthr", [(True, 0.5, 0.5), (True, -1.0, 0.5), (True, 0.5, -1.0),
                                             (False, 0.5, -1.0), (False, 0.0, -1.0)])
def test_skip_flags_match_definition(oracle, masked, thr, mthr):
    """Per-IA skip flags, with IA maxima placed exactly at the thresholds (strict <)."""
    r = np.random.default_rng(31)
    a = r.uniform(0.0, 0.45, (40, 36))
    m = r.uniform(0.0, 0.45, (40, 36))
    a[:, 18:] += 0.5  # image 1 passes the threshold on the right, the mask at the bottom
    m[20:, :] += 0.5
    a[3, 4] = 0.5     # the first IA of image 1 reaches the threshold exactly: kept by it
    m[5, 25] = 0.5    # a top-right IA of the mask reaches it exactly
    a[24:, :10] -= 1.0  # negative maxima: skipped at threshold 0 too
    p = O.make_params("mask_nsqecc" if masked else "nsqecc", 2, (8, 6), (2, 2), (5, 4), True,
                      np.float64, thr, mthr)
    got = O.skip_flags(a, p, mask=m if masked else None)
    ref = _skip_def(a, m, p, masked)
    np.testing.assert_array_equal(got, ref)
    assert got.any() and not got.all()


def test_two_mask_driver_reads_mask_vf_for_skips(oracle):
    """The two-mask restatement: flags from mask_VF change which IAs run; with
    mask_VF = mask_IA it equals the one-mask call."""
    a, b = _pair((48, 44), 41)
    m_ia = _binary_mask(a.shape, 42)
    m_vf = np.ones_like(m_ia)
    m_vf[:20, :] = 0.0
    p = O.make_params("mask_nsqecc", 2, (12, 10), (5, 4), (8, 7), True, np.float64, -1.0, 0.5)
    one, sn1, _ = O.piv(a, b, p, mask=m_ia)
    same, sn2, _ = O.piv(a, b, p, mask=m_ia, skip_mask=O.skip_flags(a, p, mask=m_ia))
    np.testing.assert_array_equal(one, same)
    fl = O.skip_flags(a, p, mask=m_vf)
    two, _, _ = O.piv(a, b, p, mask=m_ia, skip_mask=fl)
    vf = O.vf_size(p, a.shape)
    skipped = fl.reshape(vf, order="F").astype(bool)
    assert skipped.any() and (~skipped).any()
    assert np.all(two[:, skipped] == 0)
    np.testing.assert_array_equal(two[:, ~skipped], one[:, ~skipped])
