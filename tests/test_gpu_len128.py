"""FP64 transforms at search size 128 (interSize 64, searchMargin 32: the
IA-64 pairs of the batched masked workload and ZNCC with SN at that size).
Their 128 x 128 transform buffer does not fit in LDS, so the kernel keeps it in
a per-block global buffer and runs as a persistent grid; these cases check it
against the Float64 oracle at the full bar (tests/parity.py), including frames
with many more IAs than resident blocks."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from multi_quickpiv_amd import PIV, setPIVParameters, synth
from tests.parity import compare

pytestmark = pytest.mark.gpu


def _pair(shape, d, seed):
    n = synth.n_particles(shape, 0.05)
    a, b = synth.particle_pair(shape, n, 1.2, synth.uniform(d), seed, dtype=torch.float32)
    return a.numpy(), b.numpy()


def _check(oracle, a, b, alg, isz, sm, st, mask=None, mode="auto", omode=None):
    p = setPIVParameters(corr_alg="nsqecc" if alg == "mask_nsqecc" else alg, interSize=isz, searchMargin=sm,
                         step=st, computeSN=True)
    print(f"case alg={alg} mode={mode} shape={a.shape} is={isz} sm={sm} st={st}", flush=True)
    args = (mask, p) if mask is not None else (p,)
    VF, SN, PK = PIV(a, b, *args, return_peaks=True, mode=mode)
    op = oracle.make_params(alg, 2, isz, sm, st, True, a.dtype)
    VFo, SNo, aux = oracle.piv(a, b, op, mask=mask, mode=oracle.DIRECT if omode is None else omode,
                               nthreads=16, want_aux=True)
    rep = compare(VF, SN, PK, VFo, SNo, aux)
    print(rep, flush=True)
    return rep


@pytest.mark.parametrize("alg", ["mask_nsqecc", "zncc", "nsqecc", "fft"])
def test_search_size_128_fp64(oracle, alg):
    a, b = _pair((320, 288), (3.3, -4.6), 41)
    mask = synth.ellipse_mask(a.shape, n=6, cover=0.3, seed=5).numpy() if alg == "mask_nsqecc" else None
    mode = "auto" if alg == "zncc" else "fp64"  # AUTO runs ZNCC+SN at search size 128 in FP64
    rep = _check(oracle, a, b, alg, 64, 32, 32, mask=mask, mode=mode)
    assert rep.ok(), rep
    assert rep.vf_maxerr <= 1e-6, rep
    if alg == "mask_nsqecc":
        # AUTO: FP32 transforms with the IAs they flag recomputed at length 128
        # in FP64 (global transform buffers, persistent list launch)
        rep = _check(oracle, a, b, alg, 64, 32, 32, mask=mask, mode="auto")
        assert rep.ok(), rep


def test_anisotropic_window_at_128(oracle):
    """Search size 128 along dim 1 only (64 + 2*32), 96 along dim 2."""
    a, b = _pair((256, 300), (-2.2, 5.1), 43)
    rep = _check(oracle, a, b, "zncc", (64, 48), (32, 24), (24, 30), mode="fp64")
    assert rep.ok(), rep


@pytest.mark.parametrize("mode", ["fp64", "auto"])
def test_persistent_grid_many_ias(oracle, mode):
    """2,745 IAs (more than one pass of the resident blocks) of masked NSQECC,
    checked against the oracle's FFT mode (pinned to its direct mode); in AUTO
    the FP64 recomputation of the flagged IAs runs a persistent grid over the
    refine list with the global transform buffers."""
    a, b = _pair((1024, 768), (1.7, -2.9), 47)
    mask = synth.ellipse_mask(a.shape, n=20, cover=0.3, seed=9).numpy()
    rep = _check(oracle, a, b, "mask_nsqecc", 64, 32, 16, mask=mask, mode=mode, omode=oracle.FFT)
    assert rep.n == 61 * 45
    assert rep.ok(), rep
