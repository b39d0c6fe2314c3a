"""The allocation-free entry point (mqpiv_piv_device_noalloc) through the C ABI
against mqpiv_piv_device on the same device buffers: bit-identical vector
fields and SN, on two calls in a row (the second with a smaller input and a
different measure, so the shared static scratch is reused), plus the host-side
rejections (integral arrays larger than the scratch, 3-D)."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest
import torch

from multi_quickpiv_amd import _lib, setPIVParameters, synth
from multi_quickpiv_amd.piv import to_c_params

pytestmark = pytest.mark.gpu


def _call(ctx, fn, p, a, b, mode="auto"):
    shape = tuple(a.shape)
    p.ndims = len(shape)  # as PIV sets it from its inputs
    cp = to_c_params(p, None, mode)
    dims = (C.c_int64 * 3)(*(list(shape) + [1] * (3 - len(shape))))
    n = [(s - i) // st + 1 for s, i, st in zip(shape, p.interSize[: len(shape)], p.step[: len(shape)])]
    VF = torch.full((len(shape), *n), 7.0, dtype=torch.float64, device="cuda")
    SN = torch.full(tuple(n), 7.0, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()  # the library's stream is not ordered after torch's
    rc = fn(ctx.handle, C.byref(cp), dims, _lib.MQPIV_F32, a.data_ptr(), b.data_ptr(), None, None, _lib.MQPIV_F64,
            VF.data_ptr(), SN.data_ptr(), None)
    return rc, VF, SN


def _pair(shape, seed):
    a, b = synth.particle_pair(shape, synth.n_particles(shape, 0.05), 1.2, synth.uniform((2.4, -1.7)), seed,
                               device="cuda", dtype=torch.float32)
    return a.contiguous(), b.contiguous()


def test_noalloc_equals_device_entry_point_twice():
    ctx = _lib.context(0)
    cases = [
        # FP32-transform ZNCC: two 2048^2 FP64 integral arrays, ~67 MB of the scratch
        ((2048, 2048), "zncc", 32, 16, 16, 11, "fp32"),
        # smaller input, one array
        ((384, 512), "nsqecc", 24, 12, 12, 12, "auto"),
        # FP64 ZNCC forms its window sums in the kernel: no scratch
        ((256, 320), "zncc", 32, 16, 16, 13, "auto"),
    ]
    for shape, alg, isz, sm, st, seed, mode in cases:
        print(f"case noalloc alg={alg} mode={mode} shape={shape} is={isz} sm={sm} st={st}", flush=True)
        a, b = _pair(shape, seed)
        p = setPIVParameters(corr_alg=alg, interSize=isz, searchMargin=sm, step=st, computeSN=True)
        torch.cuda.synchronize()
        rc1, VF1, SN1 = _call(ctx, ctx.lib.mqpiv_piv_device, p, a, b, mode)
        ctx.check(rc1)
        torch.cuda.synchronize()
        rc2, VF2, SN2 = _call(ctx, ctx.lib.mqpiv_piv_device_noalloc, p, a, b, mode)
        ctx.check(rc2)
        torch.cuda.synchronize()
        v1, v2 = VF1.cpu().numpy(), VF2.cpu().numpy()
        s1, s2 = SN1.cpu().numpy(), SN2.cpu().numpy()
        assert np.all(np.isfinite(v1)) and not np.any(v1 == 7.0)
        assert np.array_equal(v1, v2, equal_nan=True)
        assert np.array_equal(s1, s2, equal_nan=True)


def test_noalloc_rejects_what_it_cannot_hold():
    """Host-side rejections, before anything is enqueued.  The buffers are
    allocated at the sizes the calls name, so that a call that were not
    rejected would still stay in bounds."""
    ctx = _lib.context(0)
    # NSQECC on a 4096^2 pair: its FP64 integral array (134 MB) exceeds the 96 MiB scratch
    p = setPIVParameters(corr_alg="nsqecc", interSize=32, searchMargin=16, step=16, computeSN=True)
    p.ndims = 2
    cp = to_c_params(p, None, "auto")
    big = (C.c_int64 * 3)(4096, 4096, 1)
    img = torch.zeros(4096 * 4096, dtype=torch.float32, device="cuda")
    n = ((4096 - 32) // 16 + 1) ** 2
    VF = torch.zeros(2 * n, dtype=torch.float64, device="cuda")
    SN = torch.zeros(n, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    rc = ctx.lib.mqpiv_piv_device_noalloc(ctx.handle, C.byref(cp), big, _lib.MQPIV_F32, img.data_ptr(),
                                          img.data_ptr(), None, None, _lib.MQPIV_F64, VF.data_ptr(),
                                          SN.data_ptr(), None)
    assert rc == _lib.MQPIV_ERR_OOM
    assert b"static scratch" in ctx.lib.mqpiv_last_error(ctx.handle)
    # 3-D: not taken (it needs transform scratch)
    p3 = setPIVParameters(corr_alg="nsqecc", interSize=16, searchMargin=4, step=16)
    p3.ndims = 3
    cp3 = to_c_params(p3, None, "auto")
    d3 = (C.c_int64 * 3)(32, 32, 32)
    vol = torch.zeros(32 ** 3, dtype=torch.float32, device="cuda")
    VF3 = torch.zeros(3 * 8, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    rc = ctx.lib.mqpiv_piv_device_noalloc(ctx.handle, C.byref(cp3), d3, _lib.MQPIV_F32, vol.data_ptr(),
                                          vol.data_ptr(), None, None, _lib.MQPIV_F64, VF3.data_ptr(), None, None)
    assert rc == _lib.MQPIV_ERR_UNSUPPORTED
