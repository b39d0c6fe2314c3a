"""ZNCC with SN in precision AUTO: the split-precision matrix-core form, then FP64
transforms over the IAs it flags (mqpiv_transform_form form 3).

Each IA's outputs come from exactly one of the two forms: a recomputed IA
equals the FP64 mode's result bit for bit, any other IA equals the split form's
("mfma" mode) bit for bit.  Against the FP64 mode's results (which the parity
tests hold to the Float64 oracle at 1e-13) the mix keeps identical integer
peaks, SN within 1e-4 relative and displacements within 0.01 px
(tests/parity.py), in single calls, batches and graph replays."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from multi_quickpiv_amd import setPIVParameters, synth
from multi_quickpiv_amd.piv import BatchGraph, DeviceFrame, DeviceRequest, piv_batch, transform_form

pytestmark = pytest.mark.gpu


def _frame(shape, seed, bg=0.1, noise=0.02, density=0.05):
    n = synth.n_particles(shape, density)
    a, b = synth.particle_pair(shape, n, 1.2, synth.lamb_oseen(shape, core_radius=shape[0] / 6, peak=6.0), seed,
                               bg=bg, noise=noise, device="cuda", dtype=torch.float32)
    return DeviceFrame(a, b)


def _run(frame, p, mode):
    r = DeviceRequest.from_frame(frame, p, want_peaks=True, mode=mode)
    piv_batch([r])
    torch.cuda.synchronize()
    return r.results()


def _same(got, want):
    for x, y in zip(got, want):
        assert (x is None and y is None) or np.array_equal(x, y, equal_nan=True)


def _split(VF, SN, PK, ref):
    """Per IA: outputs differ from `ref` (VF, SN, PK) anywhere."""
    VFr, SNr, PKr = ref
    n = SN.size
    d = (SN.reshape(n) != SNr.reshape(n)) & ~(np.isnan(SN.reshape(n)) & np.isnan(SNr.reshape(n)))
    d |= np.any(VF.reshape(2, n) != VFr.reshape(2, n), axis=0)
    d |= np.any(PK.reshape(2, n) != PKr.reshape(2, n), axis=0)
    return d


@pytest.mark.parametrize("isz", [32, 64])
@pytest.mark.parametrize("kind", ["plain", "background"])
def test_each_ia_from_one_form_and_at_the_bar(isz, kind):
    shape = (768, 704)
    frame = _frame(shape, 900 + isz, **({"bg": 10.0, "noise": 0.5} if kind == "background" else {}))
    p = setPIVParameters(corr_alg="zncc", interSize=isz, searchMargin=16, step=16, computeSN=True)
    assert transform_form(p, shape, "auto")[0] == "split-f16+f64"
    auto = _run(frame, p, "auto")
    split = _run(frame, p, "mfma")
    f64 = _run(frame, p, "fp64")
    refined = _split(*auto, split)
    n = refined.size
    print(f"isz={isz} {kind}: {int(refined.sum())} of {n} IAs recomputed", flush=True)
    assert 0 < refined.sum() < n // 2
    # recomputed IAs: the FP64 results; the others: the split form's
    assert not _split(*auto, f64)[refined].any()
    VF, SN, PK = auto
    VF6, SN6, PK6 = f64
    assert np.array_equal(PK, PK6)
    rel = np.abs(SN.reshape(n) / SN6.reshape(n) - 1)
    assert np.nanmax(rel) <= 1e-4
    assert np.nanmax(np.abs(VF - VF6)) <= 0.01


def test_batch_and_graph_replays_equal_single_calls():
    """ZNCC runs with different FP64 lengths (search sizes 64 and 48) and one
    without SN, in one batch and in graph replays: each equals its own call."""
    frame = _frame((384, 448), 77)
    specs = [(32, 16, True), (24, 12, True), (32, 16, False), (64, 16, True)]
    reqs, singles = [], []
    for isz, sm, sn in specs:
        p = setPIVParameters(corr_alg="zncc", interSize=isz, searchMargin=sm, step=16, computeSN=sn)
        reqs.append(DeviceRequest.from_frame(frame, p, want_peaks=True))
        singles.append(_run(frame, p, "auto"))
    piv_batch(reqs)
    torch.cuda.synchronize()
    for r, one in zip(reqs, singles):
        _same(r.results(), one)
    g = BatchGraph(reqs)
    for _ in range(3):
        for r in reqs:
            r.VF.fill_(3.0)
            if r.SN is not None:
                r.SN.fill_(3.0)
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        for r, one in zip(reqs, singles):
            _same(r.results(), one)
    g.close()


def _run_masked(a, b, m, p, mode):
    r = DeviceRequest(a, b, p, mask=m, mode=mode, want_peaks=True)
    piv_batch([r])
    torch.cuda.synchronize()
    return r.results()


@pytest.mark.parametrize("isz", [16, 32, 64])
def test_masked_each_ia_from_one_form_and_at_the_bar(isz):
    """Masked NSQECC in AUTO (form 4): FP32 transforms, the IAs they flag
    recomputed by the FP64 transforms of the same length."""
    shape = (768, 704)
    n = synth.n_particles(shape, 0.05)
    a, b = synth.particle_pair(shape, n, 1.2, synth.uniform((2.7, -4.1)), 950 + isz, device="cuda",
                               dtype=torch.float32)
    m = synth.ellipse_mask(shape, n=20, cover=0.3, seed=960 + isz, device="cuda")
    p = setPIVParameters(corr_alg="nsqecc", interSize=isz, searchMargin=isz // 2, step=isz // 2, computeSN=True)
    assert transform_form(p, shape, "auto", masked=True)[0] == "f32+f64"
    auto = _run_masked(a, b, m, p, "auto")
    f32 = _run_masked(a, b, m, p, "fp32")
    f64 = _run_masked(a, b, m, p, "fp64")
    refined = _split(*auto, f32)
    nia = refined.size
    print(f"masked isz={isz}: {int(refined.sum())} of {nia} IAs recomputed", flush=True)
    assert refined.sum() < nia // 4
    assert not _split(*auto, f64)[refined].any()
    VF, SN, PK = auto
    VF6, SN6, PK6 = f64
    assert np.array_equal(PK, PK6)
    both_nan = np.isnan(SN) & np.isnan(SN6)
    rel = np.where(both_nan, 0.0, np.abs(SN / SN6 - 1))
    assert np.nanmax(rel) <= 1e-4
    err = np.where(np.isnan(VF) & np.isnan(VF6), 0.0, np.abs(VF - VF6))
    assert np.max(err) <= 0.01
