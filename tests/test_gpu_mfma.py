"""GPU parity of the split-precision matrix-core form (mode "mfma") against the
Float64 oracle (bar: tests/parity.py), and the calls it refuses.

The form computes the same fully-overlapping correlation as the FFT kernels
(fftcc.jl:101-126) as chained products with DFT matrices on hi + lo FP16
operands accumulated in FP32, so its error is of the order of an FP32
transform's."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from multi_quickpiv_amd import PIV, setPIVParameters, synth
from tests.parity import compare

pytestmark = pytest.mark.gpu

CASES = [
    # shape, interSize, searchMargin, step
    ((256, 256), 32, 16, 16),       # the 2048^2 workload's IA 32 geometry (length 64)
    ((203, 187), (24, 20), (10, 6), (9, 11)),
    ((96, 112), 16, 8, 8),          # search size 32, run at length 64
    ((150, 130), (40, 30), (12, 14), (20, 15)),  # 2 x 1 sub-blocks
    ((320, 288), 64, 16, 16),       # the 2048^2 workload's IA 64 geometry (2 x 2 sub-blocks)
    ((233, 251), (48, 56), (16, 12), (19, 23)),
]


def _pair(shape, seed, d=(3.4, -2.7)):
    n = synth.n_particles(shape, 0.05)
    a, b = synth.particle_pair(shape, n, 1.2, synth.uniform(d), seed, dtype=torch.float32)
    return a.numpy(), b.numpy()


@pytest.mark.parametrize("alg", ["fft", "nsqecc", "zncc"])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_mfma_matches_oracle(oracle, alg, case):
    shape, isz, sm, st = CASES[case]
    a, b = _pair(shape, 40 + case)
    sn = alg != "zncc"
    p = setPIVParameters(corr_alg=alg, interSize=isz, searchMargin=sm, step=st, computeSN=sn)
    print(f"case mfma alg={alg} shape={shape} is={isz} sm={sm} st={st}", flush=True)
    VF, SN, PK = PIV(a, b, p, return_peaks=True, mode="mfma")
    op = oracle.make_params(alg, 2, isz, sm, st, sn, a.dtype)
    VFo, SNo, aux = oracle.piv(a, b, op, mode=oracle.DIRECT, nthreads=8, want_aux=True)
    rep = compare(VF, SN, PK, VFo, SNo, aux)
    print(rep, flush=True)
    assert rep.peak_mismatch_clear == 0 and rep.vf_maxerr <= 0.01, rep
    if sn:
        assert rep.sn_maxrel <= 1e-4, rep


def test_mfma_zncc_sn_error_is_reported(oracle):
    """ZNCC's SN with this form's transform error (stated, not held to the
    Float64-path bar; the FP64 modes cover it)."""
    shape, isz, sm, st = CASES[0]
    a, b = _pair(shape, 77)
    p = setPIVParameters(corr_alg="zncc", interSize=isz, searchMargin=sm, step=st, computeSN=True)
    VF, SN, PK = PIV(a, b, p, return_peaks=True, mode="mfma")
    op = oracle.make_params("zncc", 2, isz, sm, st, True, a.dtype)
    VFo, SNo, aux = oracle.piv(a, b, op, mode=oracle.DIRECT, nthreads=8, want_aux=True)
    rep = compare(VF, SN, PK, VFo, SNo, aux)
    print(rep, flush=True)
    assert rep.peak_mismatch_clear == 0 and rep.vf_maxerr <= 0.01, rep


@pytest.mark.parametrize("what", ["3d", "masked", "search_size", "multipass"])
def test_mfma_refuses_what_it_does_not_run(what):
    if what == "3d":
        a = np.zeros((40, 40, 40), np.float32)
        p = setPIVParameters(corr_alg="nsqecc", interSize=16, searchMargin=4, step=16)
        with pytest.raises(RuntimeError, match="matrix-core"):
            PIV(a, a, p, mode="mfma")
    elif what == "masked":
        a = np.zeros((96, 96), np.float32)
        p = setPIVParameters(corr_alg="nsqecc", interSize=32, searchMargin=8, step=32)
        with pytest.raises(RuntimeError, match="matrix-core"):
            PIV(a, a, np.ones_like(a), p, mode="mfma")
    elif what == "search_size":
        a = np.zeros((256, 256), np.float32)
        for isz, sm in ((64, 17), (32, 24), (96, 8)):
            p = setPIVParameters(corr_alg="nsqecc", interSize=isz, searchMargin=sm, step=32)
            with pytest.raises(RuntimeError, match="matrix-core"):
                PIV(a, a, p, mode="mfma")
    else:
        a = np.zeros((128, 128), np.float32)
        p = setPIVParameters(corr_alg="nsqecc", interSize=16, searchMargin=4, step=16, multipass=2)
        with pytest.raises(RuntimeError, match="single passes"):
            PIV(a, a, p, mode="mfma")
