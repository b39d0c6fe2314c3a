"""Pins the Float64 oracle against what does not depend on it: integer closed
forms for the geometry, scipy's direct correlation, exact invariants of the
similarity measures, analytic surfaces for the peak/subpixel/SN step, and
delta-particle images with a known integer shift.  (The reference ships no tests,
fixtures or golden vectors to pin it against.)"""
from __future__ import annotations

import numpy as np
import pytest
from scipy import signal

from oracle import oracle as O


# ---------------------------------------------------------------- geometry
def test_vf_size_closed_form(oracle):
    p = O.make_params("fft", 2, 32, 16, 16)
    assert O.vf_size(p, (256, 256)) == (15, 15)           # C1: 225 IAs
    assert O.vf_size(p, (2048, 2048)) == (127, 127)       # C2a: 16,129
    p64 = O.make_params("nsqecc", 2, 64, 16, 16)
    assert O.vf_size(p64, (2048, 2048)) == (125, 125)     # C2b: 15,625
    p3 = O.make_params("fft", 3, 32, 8, 16)
    assert O.vf_size(p3, (512, 512, 256)) == (31, 31, 15)  # C3: 14,415
    assert O.vf_size(p, (31, 100)) == (0, 5)


def test_search_offsets_quirk_q1(oracle):
    """SR_TLF_off = |min(0, TLF - (SM - 1))| (piv_ops.jl:72): 2 less than the
    true out-of-bounds count on the low border."""
    p = O.make_params("fft", 2, 32, 16, 16)
    c = O.coords(p, (256, 256), 1)  # first IA
    tlf, brb, sr_tlf, sr_brb, off_lo, off_hi = c
    assert tuple(tlf[:2]) == (1, 1) and tuple(brb[:2]) == (32, 32)
    assert tuple(sr_tlf[:2]) == (1, 1) and tuple(sr_brb[:2]) == (48, 48)
    assert tuple(off_lo[:2]) == (14, 14)  # true count is 16
    assert tuple(off_hi[:2]) == (0, 0)
    c = O.coords(p, (256, 256), 2)  # TLF = 17 = SM + 1: no clipping
    assert c[4][0] == 0 and c[2][0] == 1
    c = O.coords(p, (256, 256), 225)  # last IA: TLF = 225, BRB = 256
    assert c[0][0] == 225 and c[3][0] == 256 and c[5][0] == 16 and c[4][0] == 0
    # TLF = SM - 1 and TLF = SM: offsets 0 although 2 / 1 px fall outside
    q = O.make_params("fft", 2, 8, 10, 8)
    for v, tlf1, expect in [(2, 9, 0)]:
        cc = O.coords(q, (64, 64), v)
        assert cc[0][0] == tlf1 and cc[4][0] == expect


def _good_size_real_py(n: int) -> int:
    """Independent transliteration of fft_ops.jl:164-198 (the loop bound uses
    base 11 with a factor 5, as written in the reference)."""
    if n <= 6:
        return n
    best = 2 * n
    k = 0
    while 5 * 11 ** (k + 1) < best:
        k += 1
    for n5 in range(k + 1):
        x = 5 ** n5
        while x < n:
            x *= 2
        while True:
            if x < n:
                x *= 3
            elif x > n:
                best = min(best, x)
                if x & 1:
                    break
                x >>= 1
            else:
                return n
    return best


def test_good_size_real(oracle):
    for n in (64, 96, 48, 32, 128):  # the benchmark workloads' transform sizes map to themselves
        assert O.good_size_real(n) == n
    assert O.good_size_real(13) == 16  # not 15: the 5-power loop is cut short
    assert O.good_size_real(7) == 8
    for n in range(1, 4097):
        assert O.good_size_real(n) == _good_size_real_py(n), n


# ---------------------------------------------------------------- correlation
def _rand_pair(shape, seed):
    r = np.random.default_rng(seed)
    return r.random(shape), r.random(shape)


@pytest.mark.parametrize("ndims", [2, 3])
def test_correlation_matches_scipy(oracle, ndims):
    shape = (40, 36) if ndims == 2 else (20, 18, 22)
    isz, sm, st = (8, 6, 7) if ndims == 2 else (6, 3, 5)
    a, b = _rand_pair(shape, 1)
    p = O.make_params("fft", ndims, isz, sm, st)
    vf = O.vf_size(p, shape)
    for lin in [1, 2, int(np.prod(vf)) // 2 + 1, int(np.prod(vf))]:
        co = O.coords(p, shape, lin)
        tlf, sr_tlf, sr_brb, off_lo = co[0] - 1, co[2] - 1, co[3] - 1, co[4]
        F = a[tuple(slice(tlf[d], tlf[d] + isz) for d in range(ndims))]
        ss = isz + 2 * sm
        G = np.zeros((ss,) * ndims)
        src = b[tuple(slice(sr_tlf[d], sr_brb[d] + 1) for d in range(ndims))]
        G[tuple(slice(off_lo[d], off_lo[d] + src.shape[d]) for d in range(ndims))] = src
        ref = signal.correlate(G, F, mode="valid", method="direct")  # ref[s] = sum F[x] G[x+s]
        for mode in (O.DIRECT, O.FFT):
            got = O.ia_correlation(a, b, p, lin, mode=mode)
            np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12)


def test_nsqecc_is_one_at_perfect_match_and_bounded(oracle):
    a, _ = _rand_pair((64, 64), 2)
    d = (3, -2)
    b = np.roll(a, d, axis=(0, 1))  # b[x + d] = a[x]
    p = O.make_params("nsqecc", 2, 16, 6, 16)
    S, co = O.ia_similarity(a, b, p, 6)  # interior IA (TLF = 17, 17)
    assert co[4][0] == 0 and co[4][1] == 0
    pk = np.unravel_index(np.argmax(S), S.shape)
    assert pk == (6 + d[0], 6 + d[1])
    assert S[pk] == pytest.approx(1.0, abs=1e-12)
    assert np.all(S <= 1 + 1e-12) and np.all(S > 0)


def test_zncc_bounded_and_mask_of_ones_equals_nsqecc(oracle):
    a, b = _rand_pair((48, 48), 3)
    pz = O.make_params("zncc", 2, 12, 5, 12)
    Z, _ = O.ia_similarity(a, b, pz, 5)
    assert np.all(np.abs(Z) <= 1 + 1e-12)
    pn = O.make_params("nsqecc", 2, 12, 5, 12)
    pm = O.make_params("mask_nsqecc", 2, 12, 5, 12)
    for lin in (1, 5, 9):
        S, _ = O.ia_similarity(a, b, pn, lin)
        M, _ = O.ia_similarity(a, b, pm, lin, mask=np.ones_like(a))
        np.testing.assert_allclose(M, S, rtol=1e-10, atol=1e-12)


def test_zncc_uses_search_region_mean(oracle):
    """zncc_ops.jl:22-24: G - mean(G over the whole clipped SR), zero padding kept.
    Adding a constant to image 2 changes nothing only if the mean is SR-wide."""
    a, b = _rand_pair((48, 48), 4)
    p = O.make_params("zncc", 2, 12, 5, 12)
    Z1, _ = O.ia_similarity(a, b, p, 5)
    Z2, _ = O.ia_similarity(a, b + 10.0, p, 5)
    np.testing.assert_allclose(Z1, Z2, rtol=1e-9, atol=1e-9)
    # a window's own mean is not subtracted: the result differs from the
    # textbook per-window ZNCC
    co = O.coords(p, a.shape, 5)
    tlf = co[0] - 1
    F = a[tlf[0]:tlf[0] + 12, tlf[1]:tlf[1] + 12]
    F = F - F.mean()
    s0 = co[2] - 1
    G = b[s0[0]:s0[0] + 22, s0[1]:s0[1] + 22]
    win = G[3:15, 7:19]
    textbook = np.sum(F * (win - win.mean())) / np.sqrt(np.sum(F ** 2) * np.sum((win - win.mean()) ** 2))
    assert abs(Z1[3, 7] - textbook) > 1e-6


# ---------------------------------------------------------------- peak / subpixel / SN
def _gauss_surface(ns, mu, a=(0.4, 0.3), K=-0.3):
    """S = exp(q) + K with q a separable parabola: log(1 + S - min) = q - q_min,
    so the three-point Gaussian fit recovers mu exactly."""
    i = np.arange(1, ns[0] + 1)[:, None]
    j = np.arange(1, ns[1] + 1)[None, :]
    q = -a[0] * (i - mu[0]) ** 2 - a[1] * (j - mu[1]) ** 2
    return q, K


def test_gaussian_subpixel_recovers_parabola_vertex(oracle):
    ns = (9, 9)
    mu = (4.3, 5.8)
    q, K = _gauss_surface(ns, mu)
    pk = (4, 6)
    box = q[pk[0] - 2:pk[0] + 1, pk[1] - 2:pk[1] + 1]
    q = q - box.min()                      # min over the 3x3 box is exactly 0
    S = np.exp(q) + K                      # box minimum of S is 1 + K
    got_pk, val, disp, sn = O.peak_subpixel_sn(S, (4, 4), (0, 0), (0, 0))
    assert got_pk == pk
    assert disp[0] == pytest.approx(mu[0] - 5, abs=1e-12)   # disp = mu - (SM + 1)
    assert disp[1] == pytest.approx(mu[1] - 5, abs=1e-12)


def test_peak_tie_break_first_in_column_major_and_zero_start(oracle):
    S = np.zeros((7, 7))
    S[2, 5] = 0.9
    S[5, 1] = 0.9  # earlier in column-major order (column 1 before column 5)
    pk, val, disp, sn = O.peak_subpixel_sn(S, (3, 3), (0, 0), (0, 0))
    assert pk == (6, 2)
    S = -np.ones((7, 7))
    pk, val, disp, sn = O.peak_subpixel_sn(S, (3, 3), (0, 0), (0, 0))
    assert pk == (0, 0) and disp == (-4.0, -4.0)  # nothing beats 0: -(SM+1)


def test_peak_window_and_refinement_edges(oracle):
    S = np.zeros((9, 9))
    S[0, 4] = 5.0   # outside the window when off_lo = 1
    S[1, 4] = 1.0   # on the window edge: no refinement
    pk, val, disp, sn = O.peak_subpixel_sn(S, (4, 4), (1, 0), (0, 0))
    assert pk == (2, 5) and disp == (2 - 5.0, 0.0)


def test_sn_window_upper_bound_is_2sm(oracle):
    """compute_SN reads [1+o-, 2SM-o+] (piv_ops.jl:352-364): one less than the
    peak window on the upper side (piv_ops.jl:354)."""
    S = np.ones((5, 5))
    S[4, :] = 100.0  # row 2SM+1 = 5 is outside the SN window
    S[:, 4] = 100.0
    S[1, 1] = 3.0
    *_, sn = O.peak_subpixel_sn(S, (2, 2), (0, 0), (0, 0))
    # window rows/cols 1..4: 16 values, max 3, rest 15 ones
    assert sn == pytest.approx(3.0 / ((3.0 + 15.0 - 3.0) / 15.0))


# ---------------------------------------------------------------- end to end
def _delta_images(shape, d, seed):
    r = np.random.default_rng(seed)
    a = np.zeros(shape)
    idx = tuple(r.integers(4, s - 4, 60) for s in shape)
    a[idx] = r.uniform(0.5, 1.0, 60)
    a += 0.01 * r.random(shape)
    b = np.roll(a, d, axis=tuple(range(len(shape))))
    return a, b


def test_integer_shift_expected_peaks_including_border_bias(oracle):
    """Interior IAs peak at SM+1+d; the first row/column (TLF = 1 <= SM-2) peak
    at SM-1+d because the search region is placed 2 px early (piv_ops.jl:72)."""
    d = (3, 2)
    a, b = _delta_images((96, 96), d, 5)
    p = O.make_params("nsqecc", 2, 16, 8, 16)
    VF, SN, aux = O.piv(a, b, p, want_aux=True)
    pk = aux[:2]
    assert pk[0, 2, 2] == 8 + 1 + d[0] and pk[1, 2, 2] == 8 + 1 + d[1]
    assert pk[0, 0, 2] == 8 - 1 + d[0] and pk[1, 2, 0] == 8 - 1 + d[1]
    assert VF[0, 2, 2] == pytest.approx(d[0], abs=0.05)
    assert VF[0, 0, 2] == pytest.approx(d[0] - 2, abs=0.05)
    # a shift of -2 on the first column would put the biased peak at SM-3 = 5,
    # left of the peak window [1 + off_lo, ...] = [7, ...]: unreachable there.
    co = O.coords(p, (96, 96), 1)
    assert 1 + co[4][1] == 7


def test_direct_and_fft_modes_agree_all_algorithms(oracle):
    a, b = _rand_pair((64, 56), 6)
    m = (np.random.default_rng(7).random((64, 56)) > 0.3).astype(np.float64)
    for alg in ("fft", "zncc", "nsqecc", "mask_nsqecc"):
        p = O.make_params(alg, 2, 12, 5, 9)
        V1, S1, _ = O.piv(a, b, p, mask=m if alg == "mask_nsqecc" else None, mode=O.DIRECT)
        V2, S2, _ = O.piv(a, b, p, mask=m if alg == "mask_nsqecc" else None, mode=O.FFT)
        np.testing.assert_allclose(V1, V2, atol=1e-9)
        np.testing.assert_allclose(S1, S2, rtol=1e-9)


def test_3d_runs_and_recovers_shift(oracle):
    d = (2, -1, 1)
    a, b = _delta_images((40, 40, 40), d, 8)
    p = O.make_params("nsqecc", 3, 12, 4, 12)
    VF, SN, aux = O.piv(a, b, p, want_aux=True)
    assert aux[0, 1, 1, 1] == 4 + 1 + d[0]
    assert VF[2, 1, 1, 1] == pytest.approx(d[2], abs=0.05)
