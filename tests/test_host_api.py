"""Host-side logic and the C ABI, without a GPU: parameter mirror, exports of
every symbol include/mqpiv.h declares, error paths the host can force."""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import multi_quickpiv_amd as mq
from multi_quickpiv_amd import _lib
from multi_quickpiv_amd.params import (
    FFTCC,
    NSQECC,
    ZNCC,
    _isize,
    _ovp,
    _smarg,
    _ssize,
    _step,
    get_vectorfield_size,
    parseCorr,
    setPIVParameters,
    toDims3,
)
from multi_quickpiv_amd.piv import host_skip_mask, to_c_params

ROOT = Path(__file__).resolve().parent.parent


# ------------------------------------------------------------- parameters
def test_defaults_match_reference():
    """piv_parameters.jl:68-82."""
    p = setPIVParameters()
    assert p.corr_alg == NSQECC()
    assert p.interSize == (32, 32, 32) and p.searchMargin == (10, 10, 10)
    assert p.overlap == (0, 0, 0) and p.step is None
    assert p.multipass == 1 and p.computeSN is False
    assert p.threshold == -1.0 and p.mask_threshold == -1.0
    assert p.ndims == 3 and p.ovp_th == 0.5
    assert p.unpadded and p.good_pad and not p.odd_pad


def test_parse_corr():
    assert parseCorr("FFT") == FFTCC() and parseCorr("zncc") == ZNCC()
    assert parseCorr("NSQECC") == NSQECC() and parseCorr("anything") == NSQECC()
    for bad in ("znc", "fftc", "nsqec"):
        with pytest.raises(NameError):
            parseCorr(bad)


def test_unpadded_good_pad_swap():
    """The outer constructor passes good_pad into the unpadded slot (piv_parameters.jl:58-59)."""
    p = setPIVParameters(unpadded=True, good_pad=False)
    assert p.unpadded is False and p.good_pad is True


def test_accessors_and_dims():
    assert toDims3(5) == (5, 5, 5) and toDims3((3, 4)) == (3, 4, 0) and toDims3(None) is None
    p = setPIVParameters(interSize=(32, 16), searchMargin=(8, 4), overlap=(16, 8))
    p.ndims = 2
    assert _isize(p) == (32, 16) and _smarg(p) == (8, 4) and _ovp(p) == (16, 8)
    assert _ssize(p) == (48, 24)
    assert _step(p) == (16, 8)  # step = interSize - overlap when step is nothing
    q = setPIVParameters(interSize=32, step=(10, 12))
    q.ndims = 2
    assert _step(q) == (10, 12)
    assert get_vectorfield_size((256, 256), setPIVParameters(interSize=32, step=16)) == (15, 15)


def test_c_params_resolution():
    p = setPIVParameters(corr_alg="zncc", interSize=32, searchMargin=16, overlap=16, computeSN=True,
                         threshold=0.5)
    p.ndims = 2
    cp = to_c_params(p)
    assert cp.corr_alg == 1 and cp.ndims == 2
    assert tuple(cp.inter_size) == (32, 32, 1) and tuple(cp.step) == (16, 16, 1)
    assert cp.compute_sn == 1 and cp.threshold == 0.5 and cp.precision_mode == _lib.MQPIV_AUTO
    with pytest.raises(NotImplementedError):
        bad = setPIVParameters(multipass=2)
        bad.ndims = 2
        to_c_params(bad)
    with pytest.raises(NotImplementedError):
        bad = setPIVParameters(unpadded=False)
        bad.ndims = 2
        to_c_params(bad)


def test_size_mismatch_is_assertion_error_like_reference():
    with pytest.raises(AssertionError, match="PIV inputs need to have the same size."):
        mq.PIV(np.zeros((64, 64)), np.zeros((64, 65)))
    with pytest.raises(AssertionError):
        mq.PIV(np.zeros((64, 64)), np.zeros((64, 64)), np.zeros((64, 63)))


def test_host_skip_mask_custom_filter():
    img = np.zeros((64, 64))
    img[:32, :32] = 1.0
    p = setPIVParameters(interSize=32, step=32, threshold=0.5, filtFun=lambda w: float(np.mean(w)))
    p.ndims = 2
    sk = host_skip_mask(img, None, p, masked=False)
    assert sk.tolist() == [0, 1, 1, 1]  # column-major IA order, only the first IA is bright
    m = np.ones_like(img)
    m[:, 32:] = 0
    pm = setPIVParameters(interSize=32, step=32, mask_threshold=0.5)
    pm.ndims = 2
    assert host_skip_mask(img, m, pm, masked=True).tolist() == [0, 0, 1, 1]


# ------------------------------------------------------------- C ABI
def _declared_symbols():
    text = (ROOT / "include" / "mqpiv.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(mqpiv_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = _declared_symbols()
    assert len(names) >= 9
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) <= set(_lib.EXPORTS)


def test_vf_size_and_param_errors_without_gpu():
    lib = _lib.load()
    cp = _lib.MqpivParams()
    cp.corr_alg, cp.ndims = 2, 2
    cp.inter_size[:] = [32, 32, 1]
    cp.search_margin[:] = [16, 16, 0]
    cp.step[:] = [16, 16, 1]
    cp.precision_mode = _lib.MQPIV_AUTO
    dims = (C.c_int64 * 3)(2048, 2048, 1)
    vf = (C.c_int64 * 3)()
    n = C.c_int64()
    assert lib.mqpiv_vf_size(C.byref(cp), dims, vf, C.byref(n)) == _lib.MQPIV_OK
    assert tuple(vf) == (127, 127, 1) and n.value == 16129
    cp.ndims = 4
    assert lib.mqpiv_vf_size(C.byref(cp), dims, vf, C.byref(n)) == _lib.MQPIV_ERR_BAD_PARAMS
    cp.ndims, cp.corr_alg = 2, 9
    assert lib.mqpiv_vf_size(C.byref(cp), dims, vf, C.byref(n)) == _lib.MQPIV_ERR_BAD_PARAMS
    cp.corr_alg, cp.precision_mode = 2, 5  # no such mode
    assert lib.mqpiv_vf_size(C.byref(cp), dims, vf, C.byref(n)) == _lib.MQPIV_ERR_UNSUPPORTED
    # the split-precision matrix-core form: 2-D FFTCC/ZNCC/NSQECC only
    cp.precision_mode = _lib.MQPIV_SPLIT_MFMA
    assert lib.mqpiv_vf_size(C.byref(cp), dims, vf, C.byref(n)) == _lib.MQPIV_OK
    cp.corr_alg = 3
    assert lib.mqpiv_vf_size(C.byref(cp), dims, vf, C.byref(n)) == _lib.MQPIV_ERR_UNSUPPORTED
    cp.corr_alg, cp.ndims = 2, 3
    cp.inter_size[2], cp.step[2] = 1, 1
    assert lib.mqpiv_vf_size(C.byref(cp), dims, vf, C.byref(n)) == _lib.MQPIV_ERR_UNSUPPORTED
    cp.ndims = 2
    cp.precision_mode = _lib.MQPIV_AUTO
    cp.step[0] = 0
    assert lib.mqpiv_vf_size(C.byref(cp), dims, vf, C.byref(n)) == _lib.MQPIV_ERR_BAD_PARAMS
    assert lib.mqpiv_vf_size(None, dims, vf, C.byref(n)) == _lib.MQPIV_ERR_NULL


def test_transform_form_follows_the_precision_mode():
    """mqpiv_transform_form (host only): AUTO runs masked NSQECC in FP32 transforms
    with FP64 recomputation of the IAs they flag, runs
    ZNCC with SN in the split-precision form with FP64 recomputation of the IAs
    it flags where that form takes the job (FP64 beyond), the split-precision
    form elsewhere where it takes the job (interSize <= 32 with search size
    <= 64, or <= 64 with margin <= 16), FP32 transforms beyond; explicit modes
    are honoured or refused."""
    from multi_quickpiv_amd.piv import transform_form

    def form(alg, isz, sm, sn=True, mode="auto", masked=False, shape=(2048, 2048)):
        p = setPIVParameters(corr_alg=alg, interSize=isz, searchMargin=sm, step=16, computeSN=sn)
        return transform_form(p, shape, mode, masked=masked)

    assert form("zncc", 32, 16) == ("split-f16+f64", 64)
    assert form("zncc", 64, 16) == ("split-f16+f64", 65)
    assert form("zncc", 64, 32) == ("f64", 128)  # margin 32: beyond the split form
    assert form("zncc", 32, 16, mode="mfma") == ("split-f16", 64)
    assert form("nsqecc", 32, 16) == ("split-f16", 64)
    assert form("nsqecc", 64, 16) == ("split-f16", 65)  # 2 x 2 sub-blocks
    assert form("zncc", 32, 16, sn=False) == ("split-f16", 64)
    assert form("fft", 16, 8) == ("split-f16", 64)
    assert form("nsqecc", 64, 32) == ("f32", 128)  # margin 32: beyond the split form
    assert form("nsqecc", 32, 16, masked=True) == ("f32+f64", 64)  # FP32, flagged IAs in FP64
    assert form("nsqecc", 64, 32, masked=True) == ("f32+f64", 128)
    assert form("nsqecc", 64, 32, masked=True, mode="fp64") == ("f64", 128)
    assert form("nsqecc", 32, 16, mode="fp32") == ("f32", 64)
    assert form("nsqecc", 32, 16, mode="fp64") == ("f64", 64)
    assert form("nsqecc", 64, 16, mode="mfma") == ("split-f16", 65)
    assert form("nsqecc", 32, 16, shape=(64, 64, 64)) == ("f32", 64)  # 3-D: FP32 transforms
    assert form("nsqecc", 16, 8, masked=True, shape=(64, 64, 64)) == ("f64", 32)  # 3-D masked: FP64
    with pytest.raises(RuntimeError):
        form("nsqecc", 64, 32, mode="mfma")
    with pytest.raises(RuntimeError):
        form("nsqecc", 32, 16, mode="mfma", masked=True)


def test_null_context_is_rejected():
    lib = _lib.load()
    cp = _lib.MqpivParams()
    dims = (C.c_int64 * 3)(64, 64, 1)
    assert lib.mqpiv_piv(None, C.byref(cp), dims, 0, None, None, None, None, 0, None, None) == _lib.MQPIV_ERR_NULL
    assert lib.mqpiv_ctx_destroy(None) == _lib.MQPIV_ERR_NULL
    assert lib.mqpiv_profile_begin(None) == _lib.MQPIV_ERR_NULL
    assert lib.mqpiv_version().startswith(b"mqpiv")


def test_server_argument_errors_without_gpu():
    lib = _lib.load()
    h = C.c_void_p()
    assert lib.mqpiv_server_create(0, 1 << 30, 0, 8, C.byref(h)) == _lib.MQPIV_ERR_BAD_PARAMS
    assert lib.mqpiv_server_create(0, 1 << 30, 2, 0, C.byref(h)) == _lib.MQPIV_ERR_BAD_PARAMS
    assert lib.mqpiv_server_create(0, 1024, 2, 8, C.byref(h)) == _lib.MQPIV_ERR_BAD_PARAMS
    assert lib.mqpiv_server_create(0, 1 << 30, 2, 8, None) == _lib.MQPIV_ERR_NULL
    t = C.c_uint64()
    st = C.c_int32()
    assert lib.mqpiv_server_submit(None, None, C.byref(t)) == _lib.MQPIV_ERR_NULL
    assert lib.mqpiv_server_wait(None, 1, 0.0, C.byref(st)) == _lib.MQPIV_ERR_NULL
    assert lib.mqpiv_server_stats_get(None, None) == _lib.MQPIV_ERR_NULL
    assert lib.mqpiv_server_destroy(None) == _lib.MQPIV_ERR_NULL


def test_sizes_beyond_32bit_indices_are_rejected_without_gpu():
    """The kernels index IAs and image coordinates in 32 bits; the parameter
    check (shared by every entry point) rejects larger calls up front."""
    lib = _lib.load()
    p = setPIVParameters(interSize=2, searchMargin=0, step=1)
    p.ndims = 2
    cp = to_c_params(p)
    vf = (C.c_int64 * 3)()
    n = C.c_int64()
    ok = (C.c_int64 * 3)(46340, 46341, 1)  # (46339 * 46340) IAs < 2^31
    assert lib.mqpiv_vf_size(C.byref(cp), ok, vf, C.byref(n)) == _lib.MQPIV_OK
    assert n.value == 46339 * 46340 < 2 ** 31
    many = (C.c_int64 * 3)(46342, 46342, 1)  # 46341^2 IAs > 2^31
    assert lib.mqpiv_vf_size(C.byref(cp), many, vf, C.byref(n)) == _lib.MQPIV_ERR_UNSUPPORTED
    wide = (C.c_int64 * 3)(2 ** 31, 4, 1)
    assert lib.mqpiv_vf_size(C.byref(cp), wide, vf, C.byref(n)) == _lib.MQPIV_ERR_UNSUPPORTED


def test_multipass_argument_checks_without_gpu():
    """multipass > 1 takes the default filtFun, at most one mask and no peaks
    output; multipass < 1 is rejected; all before any device call."""
    img = np.random.default_rng(0).random((64, 64))
    p = setPIVParameters(interSize=16, searchMargin=4, multipass=2, filtFun=lambda w: float(np.mean(w)),
                         threshold=0.1)
    with pytest.raises(NotImplementedError):
        mq.PIV(img, img, p)
    p = setPIVParameters(interSize=16, searchMargin=4, multipass=2)
    with pytest.raises(NotImplementedError):
        mq.PIV(img, img, img, img, p)
    with pytest.raises(NotImplementedError):
        mq.PIV(img, img, p, return_peaks=True)
    p = setPIVParameters(multipass=0)
    p.ndims = 2
    with pytest.raises(ValueError):
        to_c_params(p, multipass_ok=True)


def test_batch_requests_reject_multipass_without_gpu():
    """Batches, graphs and the server run single passes; a multipass request is
    refused when it is built, before any device work."""
    import torch

    from multi_quickpiv_amd.piv import DeviceFrame, DeviceRequest

    img = torch.rand(64, 64)
    frame = DeviceFrame(img, img)
    p = setPIVParameters(interSize=16, searchMargin=4, multipass=2)
    with pytest.raises(NotImplementedError):
        DeviceRequest.from_frame(frame, p)
