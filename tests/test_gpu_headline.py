"""The benchmark's own workloads under the parity bar (tests/parity.py):

* the 2-D headline step: the FP32 2048 x 2048 Lamb-Oseen particle pair with
  ZNCC and NSQECC at interSize 32 and 64 (searchMargin 16, step 16, computeSN),
  the four runs as one batched call on one device frame, exactly as bench.py
  times it (63,508 IAs), against the Float64 oracle's FFT mode (which the CPU
  tests pin to its direct mode at 1e-9);
* the 3-D volume workload (512 x 512 x 256, interSize 32^3, searchMargin 8,
  step 16): all 14,415 IAs on the device, checked on ranges of IAs (the first
  ones, which include the clipped border regions, and interior ones) that the
  oracle finishes in seconds."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from multi_quickpiv_amd import setPIVParameters, synth
from multi_quickpiv_amd.piv import DeviceFrame, DeviceRequest, piv_batch
from tests.parity import compare

pytestmark = pytest.mark.gpu


def test_headline_step_2048_zncc_nsqecc(oracle):
    a, b = synth.config_c2(2048, device="cuda")
    frame = DeviceFrame(a, b)
    runs = []
    for alg in ("zncc", "nsqecc"):
        for isz in (32, 64):
            p = setPIVParameters(corr_alg=alg, interSize=isz, searchMargin=16, step=16, computeSN=True)
            runs.append((alg, isz, DeviceRequest.from_frame(frame, p, want_peaks=True)))
    piv_batch([r for _, _, r in runs])
    an, bn = a.cpu().numpy(), b.cpu().numpy()
    for alg, isz, r in runs:
        VF, SN, PK = r.results()
        op = oracle.make_params(alg, 2, isz, 16, 16, True, np.float32)
        VFo, SNo, aux = oracle.piv(an, bn, op, mode=oracle.FFT, nthreads=16, want_aux=True)
        rep = compare(VF, SN, PK, VFo, SNo, aux)
        print(f"headline alg={alg} is={isz}: {rep}", flush=True)
        assert rep.n == (16129 if isz == 32 else 15625)
        assert rep.ok(), rep


def _subset(VF, SN, PK, VFo, SNo, aux, b, e):
    N = VF.shape[0]
    cut = lambda x, k: x.reshape(k, -1, order="F")[:, b:e]
    return (cut(VF, N), None if SN is None else SN.reshape(-1, order="F")[b:e], cut(PK, N), cut(VFo, N),
            None if SNo is None else SNo.reshape(-1, order="F")[b:e], cut(aux, 6))


@pytest.mark.parametrize("alg,sn", [("fft", False), ("zncc", False), ("zncc", True)])
def test_volume_512x512x256_ranges(oracle, alg, sn):
    a, b = synth.config_c3(device="cuda")
    frame = DeviceFrame(a, b)
    p = setPIVParameters(corr_alg=alg, interSize=32, searchMargin=8, step=16, computeSN=sn)
    r = DeviceRequest.from_frame(frame, p, want_peaks=True)
    piv_batch([r])
    out = r.results()
    VF, PK = out[0], out[-1]
    SN = out[1] if sn else None
    del frame, r
    an, bn = a.cpu().numpy(), b.cpu().numpy()
    del a, b
    torch.cuda.empty_cache()
    n = int(np.prod(VF.shape[1:]))
    assert n == 14415
    op = oracle.make_params(alg, 3, 32, 8, 16, sn, np.float32)
    for lo, hi in ((0, 96), (7200, 7296), (n - 64, n)):
        VFo, SNo, aux = oracle.piv(an, bn, op, mode=oracle.FFT, nthreads=16, ia_begin=lo, ia_end=hi,
                                   want_aux=True)
        sub = _subset(VF, SN, PK, VFo, SNo, aux, lo, hi)
        rep = compare(*sub)
        print(f"volume alg={alg} sn={sn} IAs [{lo}, {hi}): {rep}", flush=True)
        if not rep.ok():
            print("device peaks", sub[2][:, :3].T.tolist(), "oracle", sub[5][:3, :3].T.tolist(), flush=True)
        assert rep.n == hi - lo
        assert rep.ok(), rep
