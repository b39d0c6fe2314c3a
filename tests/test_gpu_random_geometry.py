"""Seeded random geometries on the GPU against the Float64 oracle (bar:
tests/parity.py): anisotropic interrogation areas, margins and steps, ragged
image sizes, every algorithm and precision mode, so that the window placement,
clipping and normalisation are exercised away from the hand-picked cases."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from multi_quickpiv_amd import PIV, setPIVParameters, synth
from tests.parity import compare

pytestmark = pytest.mark.gpu

ALGS = ("fft", "zncc", "nsqecc", "mask_nsqecc")
MODES = ("auto", "fp64", "fp32")


def _case(i: int):
    rng = np.random.default_rng(900 + i)
    alg = ALGS[i % len(ALGS)]
    mode = MODES[(i // len(ALGS)) % len(MODES)]
    lim = 96 if mode == "fp64" or (mode == "auto" and alg in ("zncc", "mask_nsqecc")) else 128
    # (the draws keep the limit they were made with; search size 128 in FP64 is
    # covered by test_gpu_len128.py)
    while True:
        isz = tuple(int(v) for v in rng.integers(6, 41, 2))
        sm = tuple(int(v) for v in rng.integers(1, 13, 2))
        if all(a + 2 * b <= lim for a, b in zip(isz, sm)):
            break
    st = tuple(int(rng.integers(max(2, a // 3), a + 1)) for a in isz)
    shape = tuple(int(a + rng.integers(0, 5) * s + rng.integers(0, s)) + 2 * b for a, s, b in zip(isz, st, sm))
    d = tuple(float(v) for v in rng.uniform(-3, 3, 2))
    return alg, mode, isz, sm, st, shape, d


@pytest.mark.parametrize("i", range(24))
def test_random_geometry_matches_oracle(oracle, i):
    alg, mode, isz, sm, st, shape, d = _case(i)
    n = synth.n_particles(shape, 0.05)
    a, b = synth.particle_pair(shape, n, 1.2, synth.uniform(d), 500 + i, dtype=torch.float32)
    a, b = a.numpy(), b.numpy()
    mask = None
    if alg == "mask_nsqecc":
        mask = np.ones_like(a)
        mask[: shape[0] // 3, : shape[1] // 4] = 0.0
    p = setPIVParameters(corr_alg="nsqecc" if alg == "mask_nsqecc" else alg, interSize=isz, searchMargin=sm,
                         step=st, computeSN=True)
    print(f"case {i}: alg={alg} mode={mode} is={isz} sm={sm} st={st} shape={shape}", flush=True)
    args = (mask, p) if mask is not None else (p,)
    VF, SN, PK = PIV(a, b, *args, return_peaks=True, mode=mode)
    op = oracle.make_params(alg, 2, isz, sm, st, True, a.dtype)
    VFo, SNo, aux = oracle.piv(a, b, op, mask=mask, mode=oracle.DIRECT, nthreads=8, want_aux=True)
    rep = compare(VF, SN, PK, VFo, SNo, aux)
    print(rep, flush=True)
    # FP32 transforms do not keep ZNCC's SN or masked NSQECC's cancelling
    # numerator at the bar in general (AUTO recomputes such IAs in FP64), but
    # these seeded fp32 cases (9, 11, 21, 23) do: no flipped peak, SN within
    # 2.5e-6, VF within 1.5e-6 px.  They are held to the full bar, so that a
    # drift of the FP32 kernels fails here.
    assert rep.ok(), rep
