"""GPU parity of multipass PIV (mqpiv_piv_multipass / _device) against the
oracle's multipass restatement (orc_piv_multipass), in FP64 transforms: the
vector fields agree to 1e-6 px (every pass's integer shift is a rounded FP64
value, so both paths take the same shifts) and SN to the single-pass bar of
tests/parity.py (ZNCC's SN is ill-conditioned where its window sum cancels)."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest
import torch

from multi_quickpiv_amd import PIV, _lib, setPIVParameters, synth
from multi_quickpiv_amd.piv import piv_device, to_c_params

pytestmark = pytest.mark.gpu

from tests.parity import SN_RTOL  # noqa: E402

VF_TOL = 1e-6


def _pair2(shape, seed, peak=7.0):
    n = synth.n_particles(shape, 0.05)
    d = synth.lamb_oseen(shape, core_radius=min(shape) / 4, peak=peak, drift=(2.2, -3.1))
    a, b = synth.particle_pair(shape, n, 1.2, d, seed, dtype=torch.float32)
    return a.numpy(), b.numpy()


def _check(VF, SN, VFo, SNo):
    assert VF.shape == VFo.shape
    err = np.abs(VF - VFo)
    assert float(np.max(err)) <= VF_TOL, float(np.max(err))
    if SNo is not None:
        both = np.isnan(SN) & np.isnan(SNo)
        with np.errstate(invalid="ignore", divide="ignore"):
            rel = np.where(both, 0.0, np.abs(SN - SNo) / np.maximum(np.abs(SNo), 1e-300))
        assert float(np.nanmax(rel)) <= SN_RTOL and not np.isnan(rel).any()


@pytest.mark.parametrize("alg", ["fft", "zncc", "nsqecc", "mask_nsqecc"])
@pytest.mark.parametrize("m", [2, 3])
def test_2d_multipass_matches_oracle(oracle, alg, m):
    a, b = _pair2((200, 168), seed=31 + m)
    mask = None
    if alg == "mask_nsqecc":
        mask = synth.ellipse_mask(a.shape, n=6, cover=0.25, seed=4).numpy()
    p = setPIVParameters(corr_alg="nsqecc" if alg == "mask_nsqecc" else alg, interSize=16, searchMargin=4,
                         step=8, computeSN=True, multipass=m)
    print(f"case 2d alg={alg} multipass={m} shape={a.shape}", flush=True)
    args = (mask, p) if mask is not None else (p,)
    VF, SN = PIV(a, b, *args, mode="fp64")
    op = oracle.make_params(alg, 2, 16, 4, 8, True, a.dtype)
    VFo, SNo = oracle.piv_multipass(a, b, op, m, mask=mask, nthreads=8)
    _check(VF, SN, VFo, SNo)


def test_coarse_pass_without_ias_leaves_the_next_pass_unshifted(oracle):
    """40x44 image, interSize 16, three passes: pass 3 (interSize 48) has no IA,
    pass 2 (interSize 32, one IA) runs unshifted, pass 1 is shifted by pass 2."""
    a, b = _pair2((40, 44), seed=12, peak=3.0)
    p = setPIVParameters(corr_alg="nsqecc", interSize=16, searchMargin=4, step=8, computeSN=True, multipass=3)
    VF, SN = PIV(a, b, p, mode="fp64")
    op = oracle.make_params("nsqecc", 2, 16, 4, 8, True, a.dtype)
    VFo, SNo = oracle.piv_multipass(a, b, op, 3, nthreads=8)
    _check(VF, SN, VFo, SNo)


@pytest.mark.parametrize("alg", ["nsqecc", "zncc"])
def test_threshold_skips_inside_every_pass(oracle, alg):
    """A threshold that skips some IAs of each pass: skipped coarse IAs
    contribute zero vectors to the next pass's shifts, skipped final IAs stay 0."""
    a, b = _pair2((160, 152), seed=17)
    a = a.copy()
    a[:70, :60] *= 0.2  # a dim corner: its IAs fall below the threshold
    p = setPIVParameters(corr_alg=alg, interSize=16, searchMargin=4, step=8, computeSN=True, multipass=2,
                         threshold=0.5)
    VF, SN = PIV(a, b, p, mode="fp64")
    op = oracle.make_params(alg, 2, 16, 4, 8, True, a.dtype, threshold=0.5)
    VFo, SNo = oracle.piv_multipass(a, b, op, 2, nthreads=8)
    assert (VFo[:, 0, 0] == 0).all()  # the corner IA is skipped
    _check(VF, SN, VFo, SNo)


@pytest.mark.parametrize("alg", ["fft", "zncc", "nsqecc"])
def test_anisotropic_windows_on_a_ragged_image(oracle, alg):
    a, b = _pair2((203, 171), seed=44)
    p = setPIVParameters(corr_alg=alg, interSize=(16, 12), searchMargin=(4, 6), step=(8, 6), computeSN=True,
                         multipass=2)
    VF, SN = PIV(a, b, p, mode="fp64")
    op = oracle.make_params(alg, 2, (16, 12), (4, 6), (8, 6), True, a.dtype)
    VFo, SNo = oracle.piv_multipass(a, b, op, 2, nthreads=8)
    _check(VF, SN, VFo, SNo)


def test_float64_inputs_and_float32_outputs(oracle):
    a, b = _pair2((144, 160), seed=45)
    a64, b64 = a.astype(np.float64), b.astype(np.float64)
    p = setPIVParameters(corr_alg="nsqecc", interSize=16, searchMargin=4, step=8, computeSN=True, multipass=3)
    VF, SN = PIV(a64, b64, p, mode="fp64", precision=np.float32)
    assert VF.dtype == np.float32 and SN.dtype == np.float32
    op = oracle.make_params("nsqecc", 2, 16, 4, 8, True, np.float64)
    VFo, SNo = oracle.piv_multipass(a64, b64, op, 3, nthreads=8)
    assert np.abs(VF - VFo).max() <= 1e-5  # FP32 rounding of the output
    assert np.nanmax(np.abs(SN - SNo) / np.abs(SNo)) <= 1e-6


def test_fp32_transforms_follow_the_oracle_at_the_fp32_bar(oracle):
    """FP32 transforms (mode fp32): the integer shifts of each pass are rounded
    FP32 results, so a coarse value within ~1e-5 of a .5 boundary may round the
    other way; the final field agrees to the FP32 bar everywhere else."""
    a, b = _pair2((200, 168), seed=46)
    p = setPIVParameters(corr_alg="nsqecc", interSize=16, searchMargin=4, step=8, computeSN=True, multipass=2)
    VF, SN = PIV(a, b, p, mode="fp32")
    op = oracle.make_params("nsqecc", 2, 16, 4, 8, True, a.dtype)
    VFo, SNo = oracle.piv_multipass(a, b, op, 2, nthreads=8)
    err = np.abs(VF - VFo).max(axis=0)
    assert (err <= 1e-4).mean() >= 0.99, (err > 1e-4).sum()


def test_3d_three_passes_match_oracle(oracle):
    shape = (52, 44, 40)
    n = synth.n_particles(shape, 0.01)
    a, b = synth.particle_pair(shape, n, 1.5, synth.uniform((4.6, -3.7, 2.4)), 47, dtype=torch.float32)
    a, b = a.numpy(), b.numpy()
    p = setPIVParameters(corr_alg="nsqecc", interSize=8, searchMargin=1, step=4, computeSN=True, multipass=3)
    VF, SN = PIV(a, b, p, mode="fp64")
    op = oracle.make_params("nsqecc", 3, 8, 1, 4, True, a.dtype)
    VFo, SNo = oracle.piv_multipass(a, b, op, 3, nthreads=8)
    _check(VF, SN, VFo, SNo)


@pytest.mark.parametrize("alg", ["fft", "zncc"])
def test_3d_multipass_matches_oracle(oracle, alg):
    shape = (40, 36, 32)
    n = synth.n_particles(shape, 0.01)
    a, b = synth.particle_pair(shape, n, 1.5, synth.uniform((5.3, -4.6, 3.2)), 41, dtype=torch.float32)
    a, b = a.numpy(), b.numpy()
    p = setPIVParameters(corr_alg=alg, interSize=8, searchMargin=2, step=6, computeSN=True, multipass=2)
    print(f"case 3d alg={alg} multipass=2 shape={shape}", flush=True)
    VF, SN = PIV(a, b, p, mode="fp64")
    op = oracle.make_params(alg, 3, 8, 2, 6, True, a.dtype)
    VFo, SNo = oracle.piv_multipass(a, b, op, 2, nthreads=8)
    _check(VF, SN, VFo, SNo)


def test_translation_beyond_the_margin():
    """d = (7.3, -6.4) with searchMargin 3: three passes recover it on the
    interior IAs; one pass cannot."""
    D = np.array([7.3, -6.4])
    n = synth.n_particles((256, 256), 0.05)
    a, b = synth.particle_pair((256, 256), n, 1.2, synth.uniform(tuple(D)), 23, dtype=torch.float32)
    a, b = a.numpy(), b.numpy()
    p = setPIVParameters(corr_alg="nsqecc", interSize=16, searchMargin=3, step=16, multipass=3)
    VF, _ = PIV(a, b, p)
    err = np.abs(VF[:, 5:9, 5:9] - D[:, None, None])
    assert err.max() < 0.15, err.max()
    p1 = setPIVParameters(corr_alg="nsqecc", interSize=16, searchMargin=3, step=16)
    VF1, _ = PIV(a, b, p1)
    assert np.abs(VF1[:, 5:9, 5:9] - D[:, None, None]).min() > 2.3


def test_multipass_one_is_the_single_pass_call():
    a, b = _pair2((128, 112), seed=3)
    p = setPIVParameters(corr_alg="zncc", interSize=16, searchMargin=6, step=8, computeSN=True)
    VF, SN = PIV(a, b, p)
    cp = to_c_params(p)
    ctx = _lib.context(0)
    VF1 = np.zeros_like(VF, order="F")
    SN1 = np.zeros_like(SN, order="F")
    dims = (C.c_int64 * 3)(128, 112, 1)
    A, B = np.asfortranarray(a), np.asfortranarray(b)
    ctx.check(ctx.lib.mqpiv_piv_multipass(ctx.handle, C.byref(cp), 1, dims, _lib.MQPIV_F32, A.ctypes.data,
                                          B.ctypes.data, None, _lib.MQPIV_F64, VF1.ctypes.data, SN1.ctypes.data))
    np.testing.assert_array_equal(VF, VF1)
    np.testing.assert_array_equal(SN, SN1)


def test_device_entry_matches_host_entry():
    a, b = _pair2((160, 144), seed=9)
    p = setPIVParameters(corr_alg="nsqecc", interSize=16, searchMargin=4, step=8, computeSN=True, multipass=2)
    VF, SN = PIV(a, b, p)
    ctx = _lib.context(0)
    A = torch.from_numpy(a).t().contiguous().cuda()
    B = torch.from_numpy(b).t().contiguous().cuda()
    vf = VF.shape[1:]
    dVF = torch.zeros(2 * vf[0] * vf[1], dtype=torch.float64, device="cuda")
    dSN = torch.zeros(vf[0] * vf[1], dtype=torch.float64, device="cuda")
    st = torch.cuda.Stream()
    torch.cuda.synchronize()  # the buffers above were filled on torch's stream
    p.ndims = 2
    for _ in range(2):  # the second call reuses the context's buffers
        piv_device(ctx, p, a.shape, _lib.MQPIV_F32, A.data_ptr(), B.data_ptr(), None, _lib.MQPIV_F64,
                   dVF.data_ptr(), dSN.data_ptr(), st.cuda_stream)
        st.synchronize()
        np.testing.assert_array_equal(dVF.cpu().numpy().reshape((2, *vf), order="F"), VF)
        np.testing.assert_array_equal(dSN.cpu().numpy().reshape(vf, order="F"), SN)


def test_coarse_pass_beyond_the_largest_transform_is_rejected():
    a, b = _pair2((256, 256), seed=5)
    # pass 3: interSize 96 + 2 * 48 = 192 > 128
    p = setPIVParameters(corr_alg="fft", interSize=32, searchMargin=16, step=16, multipass=3)
    with pytest.raises(_lib.MqpivError) as e:
        PIV(a, b, p)
    assert e.value.code == _lib.MQPIV_ERR_UNSUPPORTED
