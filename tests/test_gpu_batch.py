"""Batched execution and graph capture: every pair's results in a mixed batch
are bit-identical to the same pair run alone, and every graph replay matches the
eager results."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from multi_quickpiv_amd import PIV, setPIVParameters, synth
from multi_quickpiv_amd.piv import BatchGraph, DeviceRequest, piv_batch

pytestmark = pytest.mark.gpu


def _pair(shape, d, seed, dtype=torch.float32):
    n = synth.n_particles(shape, 0.05)
    a, b = synth.particle_pair(shape, n, 1.2, synth.uniform(d), seed, device="cuda", dtype=dtype)
    return a, b


def _mixed_requests():
    reqs = []
    specs = [
        ((256, 224), "nsqecc", 32, 16, 16, None),
        ((192, 256), "nsqecc", 32, 12, 16, None),     # same group (L=64), other margin
        ((300, 200), "zncc", 24, 8, 12, None),
        ((256, 256), "fft", 16, 8, 8, None),
        ((384, 320), "mask_nsqecc", 32, 16, 16, "mask"),
        ((256, 288), "mask_nsqecc", 16, 8, 8, "mask"),
        ((320, 256), "nsqecc", 64, 32, 32, None),     # L=128
        ((320, 288), "mask_nsqecc", 64, 32, 32, "mask"),  # FP64 L=128: global transform buffers
        ((288, 320), "zncc", 64, 32, 16, None),       # FP64 L=128 (ZNCC with SN)
    ]
    for i, (shape, alg, isz, sm, st, mk) in enumerate(specs):
        a, b = _pair(shape, (1.5 + 0.3 * i, -2.0 + 0.5 * i), 100 + i)
        p = setPIVParameters(corr_alg=alg if alg != "mask_nsqecc" else "nsqecc", interSize=isz,
                             searchMargin=sm, step=st, computeSN=True)
        m = synth.ellipse_mask(shape, n=5, cover=0.3, seed=i, device="cuda") if mk else None
        reqs.append((a, b, p, m))
    return reqs


def test_batch_equals_single_calls():
    specs = _mixed_requests()
    dreqs = [DeviceRequest(a, b, p, mask=m, want_peaks=True) for a, b, p, m in specs]
    piv_batch(dreqs)
    torch.cuda.synchronize()
    for (a, b, p, m), r in zip(specs, dreqs):
        VFb, SNb, PKb = r.results()
        args = (m.cpu().numpy(), p) if m is not None else (p,)
        VF1, SN1, PK1 = PIV(a.cpu().numpy(), b.cpu().numpy(), *args, return_peaks=True)
        print(r.shape, np.abs(VFb - VF1).max(), flush=True)
        assert np.array_equal(PKb, PK1)
        assert np.array_equal(VFb, VF1, equal_nan=True)
        assert np.array_equal(SNb, SN1, equal_nan=True)


def test_graph_replays_match_eager():
    specs = [s for s in _mixed_requests()]
    dreqs = [DeviceRequest(a, b, p, mask=m) for a, b, p, m in specs]
    piv_batch(dreqs)
    torch.cuda.synchronize()
    eager = [r.results() for r in dreqs]
    g = BatchGraph(dreqs)
    for rep in range(3):
        for r in dreqs:
            r.VF.fill_(7.0)  # eager work between replays
            if r.SN is not None:
                r.SN.fill_(7.0)
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        for r, (VFe, SNe) in zip(dreqs, eager):
            VFg, SNg = r.results()
            assert np.array_equal(VFg, VFe, equal_nan=True), rep
            assert np.array_equal(SNg, SNe, equal_nan=True), rep
    # new frame contents in the same buffers: the replay recomputes
    a2, b2 = _pair(specs[0][0].shape, (0.5, 0.5), 999)
    dreqs[0].a.copy_(a2.t().contiguous())
    dreqs[0].b.copy_(b2.t().contiguous())
    torch.cuda.synchronize()
    g.launch()
    torch.cuda.synchronize()
    VF1, SN1 = PIV(a2.cpu().numpy(), b2.cpu().numpy(), specs[0][2])
    VFg, SNg = dreqs[0].results()
    assert np.array_equal(VFg, VF1, equal_nan=True)
    g.close()


def test_requests_on_one_frame_share_integral_arrays():
    """Requests made from one DeviceFrame read one set of integral arrays (built
    by the first request that needs the most tables); each result is still
    bit-identical to its own call, in a batch and in graph replays, including a
    frame of another size and a double-precision frame in the same batch."""
    from multi_quickpiv_amd.piv import DeviceFrame

    a, b = _pair((288, 256), (2.2, -1.4), 321)
    frame = DeviceFrame(a, b)
    other = DeviceFrame(*_pair((200, 232), (-0.7, 2.5), 322))
    a64, b64 = _pair((160, 192), (1.1, 0.4), 323, dtype=torch.float64)
    frame64 = DeviceFrame(a64, b64)
    specs = [  # NSQECC first: the ZNCC request then builds both tables itself
        (frame, (a, b), "nsqecc", 32, 16, 16),
        (frame, (a, b), "zncc", 32, 16, 16),
        (frame, (a, b), "zncc", 64, 16, 16),
        (other, None, "zncc", 24, 12, 12),
        (frame, (a, b), "nsqecc", 16, 8, 8),
        (frame64, (a64, b64), "nsqecc", 32, 16, 16),
        (frame64, (a64, b64), "zncc", 16, 8, 8),
        (frame, (a, b), "fft", 32, 16, 16),
    ]
    dreqs = []
    for fr, _, alg, isz, sm, st in specs:
        p = setPIVParameters(corr_alg=alg, interSize=isz, searchMargin=sm, step=st, computeSN=True)
        dreqs.append(DeviceRequest.from_frame(fr, p, want_peaks=True))
    assert dreqs[0].b.data_ptr() == dreqs[1].b.data_ptr()
    piv_batch(dreqs)
    torch.cuda.synchronize()
    singles = []
    for (fr, ab, alg, isz, sm, st), r in zip(specs, dreqs):
        if ab is None:
            ab = (fr.a.t().contiguous(), fr.b.t().contiguous())
        p = setPIVParameters(corr_alg=alg, interSize=isz, searchMargin=sm, step=st, computeSN=True)
        one = PIV(ab[0].cpu().numpy(), ab[1].cpu().numpy(), p, return_peaks=True)
        got = r.results()
        print(alg, isz, r.shape, flush=True)
        for x, y in zip(got, one):
            assert np.array_equal(x, y, equal_nan=True)
        singles.append(one)
    g = BatchGraph(dreqs)
    for rep in range(2):
        for r in dreqs:
            r.VF.fill_(3.0)
            r.SN.fill_(3.0)
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        for r, one in zip(dreqs, singles):
            VFg, SNg, _ = r.results()
            assert np.array_equal(VFg, one[0], equal_nan=True), rep
            assert np.array_equal(SNg, one[1], equal_nan=True), rep
    g.close()
