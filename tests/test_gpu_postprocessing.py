"""Device vector-field filters against the CPU restatement (oracle), bit for bit:
both round every operation on its own in the field's element type and reduce
boxes in the same column-major order.  The one exception is the STD mask's
global mean/std (a tree reduction on the device, sequential on the CPU): its
mask is compared away from the threshold."""
from __future__ import annotations

import numpy as np
import pytest

from multi_quickpiv_amd import PIV, setPIVParameters, synth
from multi_quickpiv_amd import postprocessing as PP
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CASES = [
    ((2, 37, 29), np.float64, (2, 2)),
    ((2, 40, 33), np.float32, (1, 3)),
    ((3, 11, 9, 7), np.float64, (1, 2, 1)),
    ((3, 8, 10, 6), np.float32, (2, 1, 1)),
]


def _field(shape, dtype, seed):
    rng = np.random.default_rng(seed)
    VF = rng.normal(size=shape)
    VF[:, ::7] *= 6.0  # outliers
    flat = VF.reshape(shape[0], -1)
    flat[:, rng.choice(flat.shape[1], size=5, replace=False)] = 0.0  # zero vectors
    flat[0, 3] = -0.0
    return np.asfortranarray(VF.astype(dtype))


def _eq(a, b):
    assert a.dtype == b.dtype and a.shape == b.shape
    assert np.array_equal(a, b, equal_nan=True), np.nanmax(np.abs(a.astype(np.float64) - b))


@pytest.mark.parametrize("shape,dtype,r", CASES)
def test_magnitudes_masks_and_replacement(shape, dtype, r):
    VF = _field(shape, dtype, 11)
    M = PP.magnitudes(VF)
    _eq(M, O.pp_magnitudes(VF))
    _eq(PP.TH_masking(1.7, VF).astype(np.uint8), O.pp_th_mask(M, 1.7, 0))
    _eq(PP.TH_masking(0.5, VF, cmp="<").astype(np.uint8), O.pp_th_mask(M, 0.5, 1))
    dmask = PP.STD_masking(1.5, VF).astype(np.uint8)
    omask, mean, sd = O.pp_std_mask(M, 1.5)
    far = np.abs(np.abs(M.astype(np.float64) - mean) - 1.5 * sd) > 1e-9 * sd
    assert np.array_equal(dmask[far], omask[far]) and omask.sum() > 0
    # replacement with the same mask: zero and median (meanReplace = median)
    for fn, how in ((PP.zeroReplace, 0), (PP.medianReplace, 1), (PP.meanReplace, 1)):
        got = PP.STD_filtering(1.5, VF, replaceFun=fn, radius=r)
        _eq(got, O.pp_replace(VF, dmask, how, r))


@pytest.mark.parametrize("shape,dtype,r", CASES)
def test_average_collectiveness_similarity_lmm(shape, dtype, r):
    VF = _field(shape, dtype, 12)
    _eq(PP.average(r, VF), O.pp_average(VF, r))
    _eq(PP.collectiveness(r, VF), O.pp_collectiveness(VF, r, PP.cosd(40)))
    _eq(PP.collectiveness(r, VF, max_angle=75), O.pp_collectiveness(VF, r, PP.cosd(75)))
    _eq(PP.similarityAveraging(r, VF), O.pp_simavg(VF, r, PP.cosd(20)))
    for norm, scale, e in ((False, True, 2), (True, False, 2), (True, True, 1), (True, True, 5)):
        got = PP.similarityAveraging(r, VF, max_angle=30, normalize=norm, scale_by_similarity=scale,
                                     similarity_exponent=e)
        _eq(got, O.pp_simavg(VF, r, PP.cosd(30), norm, scale, e))
    out, removed = PP.local_magnitude_mask(1.8, VF, radius=r)
    oout, oremoved = O.pp_lmm(VF, 1.8, r)
    _eq(out, oout)
    assert np.array_equal(removed, oremoved.astype(bool)) and oremoved.sum() > 0


def test_filters_on_a_piv_field_and_host_callables():
    """The PIV output feeds the filters directly (same layout), and the host
    paths (callable cmp / replaceFun / fun) agree with the device paths."""
    a, b = synth.config_c1()
    p = setPIVParameters(corr_alg="fft", interSize=32, searchMargin=16, step=16, computeSN=True, threshold=0)
    VF, SN = PIV(a, b, p)
    filt = PP.remove_low_SN(SN, float(np.median(SN)), VF)
    low = SN < np.median(SN)
    assert np.all(filt[:, low] == 0) and np.array_equal(filt[:, ~low], VF[:, ~low])
    dev = PP.TH_filtering(3.0, VF, replaceFun=PP.medianReplace, radius=1)
    host = PP.TH_filtering(3.0, VF, cmp=lambda x, th: x > th, replaceFun=lambda c, rr, V: PP.medianReplace(c, rr, V),
                           radius=1)
    assert np.array_equal(dev, host)
    m_dev = PP.STD_masking(1.0, VF)
    m_fun = PP.STD_masking(1.0, VF, fun=lambda V: np.sqrt((V ** 2).sum(0)))
    assert np.array_equal(m_dev, m_fun)
    _eq(PP.average(2, VF), O.pp_average(VF, 2))


def test_empty_and_tiny_fields():
    VF = np.zeros((2, 0, 5))
    assert PP.magnitudes(VF).shape == (0, 5)
    one = np.asfortranarray(np.array([[[1.0]], [[2.0]]]))
    _eq(PP.average(3, one), one)
    assert PP.collectiveness(1, one)[0, 0] == 1.0
    out, removed = PP.local_magnitude_mask(2.0, one, radius=1)  # no neighbours: 0/0, kept = false
    _eq(out, O.pp_lmm(one, 2.0, 1)[0])
