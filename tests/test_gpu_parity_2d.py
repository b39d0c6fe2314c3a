"""GPU parity of the 2-D path (HIP kernels through the C ABI) against the
Float64 oracle, at sizes the oracle finishes in seconds.  Bar: tests/parity.py."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from multi_quickpiv_amd import PIV, setPIVParameters, synth
from multi_quickpiv_amd.piv import transform_form
from tests.parity import compare

pytestmark = pytest.mark.gpu


def _pair(shape, density=0.05, sigma=1.2, d=(2.6, -1.3), seed=7, dtype=torch.float32):
    n = synth.n_particles(shape, density)
    a, b = synth.particle_pair(shape, n, sigma, synth.uniform(d), seed, dtype=dtype)
    return a.numpy(), b.numpy()


def _run(oracle, a, b, alg, isz, sm, st, mask=None, threshold=-1.0, mask_threshold=-1.0,
         precision=np.float64, mode="auto"):
    p = setPIVParameters(corr_alg=alg if alg != "mask_nsqecc" else "nsqecc", interSize=isz,
                         searchMargin=sm, step=st, computeSN=True, threshold=threshold,
                         mask_threshold=mask_threshold)
    args = (mask, p) if mask is not None else (p,)
    print(f"case alg={alg} mode={mode} shape={a.shape} dtype={a.dtype} is={isz} sm={sm} st={st}",
          flush=True)
    VF, SN, PK = PIV(a, b, *args, precision=precision, return_peaks=True, mode=mode)
    op = oracle.make_params(alg, a.ndim, isz, sm, st, True, a.dtype, threshold, mask_threshold)
    VFo, SNo, aux = oracle.piv(a, b, op, mask=mask, mode=oracle.DIRECT, nthreads=8, want_aux=True)
    rep = compare(VF, SN, PK, VFo, SNo, aux)
    print(rep, flush=True)
    return rep, VF, SN, PK, VFo, SNo, aux


def test_config_c1_fftcc_float64(oracle):
    """256x256 Float64, 2000 particles, d=(3.4,-2.7), IA 32, SM 16, step 16, threshold 0."""
    a, b = synth.config_c1()
    rep, VF, *_ = _run(oracle, a, b, "fft", 32, 16, 16, threshold=0.0)
    assert rep.n == 225
    assert rep.ok(), rep
    # interior vectors recover the imposed shift
    assert abs(np.median(VF[0, 2:-2, 2:-2]) - 3.4) < 0.1
    assert abs(np.median(VF[1, 2:-2, 2:-2]) + 2.7) < 0.1


GEOMS = [(8, 4, 4), (16, 8, 8), (24, 12, 12), (32, 16, 16), (64, 16, 32), (64, 32, 32)]


@pytest.mark.parametrize("alg", ["fft", "zncc", "nsqecc", "mask_nsqecc"])
@pytest.mark.parametrize("isz,sm,st", GEOMS)
@pytest.mark.parametrize("mode", ["auto", "fp64", "fp32"])
def test_algorithms_all_lengths(oracle, alg, isz, sm, st, mode):
    """Every transform length of both precisions.  FP32 transforms are held to the
    full bar for FFTCC and NSQECC; for ZNCC their SN is not (the window sum of
    ZNCC values nearly cancels, so SN is ill-conditioned) and for masked NSQECC
    the num = sum(F^2) + C2 - 2 C1 cancellation flips peaks -- "auto" therefore
    runs masked NSQECC in FP64 (at search size 128, IA 64 margin 32, with the
    transform buffer in global memory), and ZNCC in the split-precision form
    with the IAs it cannot settle recomputed in FP64 (FP64 throughout where
    the split form does not take the geometry)."""
    if mode == "fp32" and alg == "mask_nsqecc":
        pytest.skip("masked NSQECC is not at the parity bar with FP32 transforms (auto uses FP64)")
    shape = (192, 160) if isz < 64 else (256, 224)
    a, b = _pair(shape)
    mask = None
    if alg == "mask_nsqecc":
        mask = synth.ellipse_mask(shape, n=6, cover=0.3, seed=3).numpy()
    rep, *_ = _run(oracle, a, b, alg, isz, sm, st, mask=mask, mode=mode)
    if mode == "fp32" and alg == "zncc":
        assert rep.peak_mismatch_clear == 0 and rep.vf_maxerr <= 0.01, rep
    else:
        assert rep.ok(), rep
    p = setPIVParameters(corr_alg=alg if alg != "mask_nsqecc" else "nsqecc", interSize=isz, searchMargin=sm,
                         step=st, computeSN=True)
    form = transform_form(p, shape, mode, masked=alg == "mask_nsqecc")[0]
    if form == "f64":
        # FP64 transforms: far inside the displacement bar
        assert rep.vf_maxerr <= 1e-6, rep
    elif form == "split-f16+f64":
        # FP32-level values where the IA was not recomputed
        assert rep.vf_maxerr <= 1e-4, rep


@pytest.mark.parametrize("alg", ["fft", "zncc", "nsqecc"])
def test_anisotropic_windows_and_ragged_image(oracle, alg):
    a, b = _pair((101, 77), d=(1.2, 3.1), seed=11)
    rep, *_ = _run(oracle, a, b, alg, (16, 24), (8, 5), (7, 9))
    assert rep.ok(), rep


def test_float64_inputs_float32_outputs(oracle):
    a, b = _pair((128, 128), dtype=torch.float64, seed=5)
    rep, VF, SN, *_ = _run(oracle, a, b, "nsqecc", 32, 10, 16, precision=np.float32)
    assert VF.dtype == np.float32 and SN.dtype == np.float32
    assert rep.ok(), rep


def test_threshold_skips_like_reference(oracle):
    a, b = _pair((160, 160), density=0.004, seed=9)
    thr = float(np.median(a)) + 0.3
    rep, VF, SN, PK, VFo, SNo, aux = _run(oracle, a, b, "nsqecc", 16, 8, 16, threshold=thr)
    skipped = (VFo[0] == 0) & (VFo[1] == 0) & (SNo == 0)
    assert skipped.any() and (~skipped).any()
    assert np.all(VF[:, skipped] == 0) and np.all(SN[skipped] == 0)
    assert rep.ok(), rep


def test_masked_fully_excluded_windows_give_nan(oracle):
    """An IA whose mask is all zero gives 0/0 everywhere: no peak, -(SM+1), SN NaN."""
    a, b = _pair((128, 128), seed=13)
    mask = np.ones_like(a)
    mask[:40, :40] = 0
    rep, VF, SN, PK, VFo, SNo, aux = _run(oracle, a, b, "mask_nsqecc", 16, 8, 16, mask=mask)
    assert np.all(PK[:, 0, 0] == 0)
    assert VF[0, 0, 0] == -9 and VF[1, 0, 0] == -9
    assert np.isnan(SN[0, 0])
    assert rep.ok(), rep


def _masks(shape):
    """An ellipse IA mask and a vector-field mask that excludes a band of rows
    and fades elsewhere, so that its maxval filter skips some IAs."""
    m_ia = synth.ellipse_mask(shape, n=6, cover=0.3, seed=5).numpy()
    m_vf = np.full(shape, 0.8, dtype=np.float32)
    m_vf[: shape[0] // 3, :] = 0.2
    m_vf[:, -40:] = 0.0
    return m_ia, m_vf


def test_mask_threshold_on_device(oracle):
    """maskThreshold > 0 with one mask: the device evaluates maxval of the IA mask
    (piv_ops.jl:148-157)."""
    a, b = _pair((192, 160), seed=17)
    m_ia, _ = _masks(a.shape)
    rep, VF, SN, PK, VFo, SNo, aux = _run(oracle, a, b, "mask_nsqecc", 16, 8, 8, mask=m_ia,
                                          mask_threshold=0.5)
    skipped = (VFo[0] == 0) & (VFo[1] == 0) & (SNo == 0)
    assert skipped.any() and (~skipped).any()
    assert np.all(VF[:, skipped] == 0) and np.all(SN[skipped] == 0)
    assert rep.ok(), rep


@pytest.mark.parametrize("thr,mthr", [(-1.0, 0.5), (0.3, 0.5), (0.3, -1.0)])
def test_two_masks_match_oracle(oracle, thr, mthr):
    """PIV(img1, img2, mask_VF, mask_IA, p) (multi_quickPIV.jl:155-213): skip flags
    from image 1 and mask_VF on the device, correlation on mask_IA."""
    a, b = _pair((192, 160), seed=19)
    a[:, :24] *= 0.25  # a dim band whose maxima fall below the image threshold
    m_ia, m_vf = _masks(a.shape)
    p = setPIVParameters(corr_alg="nsqecc", interSize=16, searchMargin=8, step=8, computeSN=True,
                         threshold=thr, mask_threshold=mthr)
    print(f"case two masks thr={thr} mthr={mthr}", flush=True)
    VF, SN, PK = PIV(a, b, m_vf, m_ia, p, return_peaks=True)
    op = oracle.make_params("mask_nsqecc", 2, 16, 8, 8, True, a.dtype, thr, mthr)
    fl = oracle.skip_flags(a, op, mask=m_vf)
    VFo, SNo, aux = oracle.piv(a, b, op, mask=m_ia, skip_mask=fl, mode=oracle.DIRECT, nthreads=8,
                               want_aux=True)
    rep = compare(VF, SN, PK, VFo, SNo, aux)
    print(rep, flush=True)
    skipped = fl.reshape(VF.shape[1:], order="F").astype(bool)
    assert skipped.any() and (~skipped).any()
    assert np.all(VF[:, skipped] == 0) and np.all(SN[skipped] == 0)
    assert rep.ok(), rep


def test_two_masks_device_entry_equals_host_entry():
    """mqpiv_piv_masks_device on resident tensors gives the host entry's outputs."""
    import ctypes as C

    from multi_quickpiv_amd import _lib
    from multi_quickpiv_amd.piv import to_c_params

    a, b = _pair((192, 160), seed=23)
    m_ia, m_vf = _masks(a.shape)
    p = setPIVParameters(corr_alg="nsqecc", interSize=16, searchMargin=8, step=8, computeSN=True,
                         threshold=0.3, mask_threshold=0.5)
    VF, SN = PIV(a, b, m_vf, m_ia, p)
    dev = [torch.from_numpy(np.asfortranarray(x).ravel(order="F")).cuda() for x in (a, b, m_vf, m_ia)]
    n = int(np.prod(SN.shape))
    dVF = torch.zeros(2 * n, dtype=torch.float64, device="cuda")
    dSN = torch.zeros(n, dtype=torch.float64, device="cuda")
    from multi_quickpiv_amd.params import mask_NSQECC

    cp = to_c_params(p, mask_NSQECC.code, "auto")
    dims = (C.c_int64 * 3)(a.shape[0], a.shape[1], 1)
    ctx = _lib.context(0)
    torch.cuda.synchronize()
    rc = ctx.lib.mqpiv_piv_masks_device(ctx.handle, C.byref(cp), dims, _lib.MQPIV_F32,
                                        *[t.data_ptr() for t in dev], _lib.MQPIV_F64, dVF.data_ptr(),
                                        dSN.data_ptr(), None)
    ctx.check(rc)
    torch.cuda.synchronize()
    got_vf = dVF.cpu().numpy().reshape(VF.shape, order="F")
    got_sn = dSN.cpu().numpy().reshape(SN.shape, order="F")
    np.testing.assert_array_equal(got_vf, VF)
    np.testing.assert_array_equal(got_sn, SN)
