"""GPU parity of the 3-D path against the Float64 oracle (bar: tests/parity.py).

3-D SN: the reference's compute_SN throws for 3-D inputs (piv_ops.jl:352-363), so
the device and the oracle both use the natural extension [1+o-, 2SM-o+] per dim;
that part is parity unpinned against the reference itself."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from multi_quickpiv_amd import PIV, setPIVParameters, synth
from tests.parity import compare

pytestmark = pytest.mark.gpu


def _vol(shape, d=(2.3, -1.1, 0.7), density=0.01, seed=21):
    n = synth.n_particles(shape, density)
    a, b = synth.particle_pair(shape, n, 1.5, synth.uniform(d), seed, dtype=torch.float32)
    return a.numpy(), b.numpy()


CASES = [
    ((40, 36, 32), 8, 4, 8),
    ((48, 48, 40), 16, 8, 12),
    ((64, 64, 64), 32, 8, 16),
    ((44, 40, 36), (8, 12, 8), (4, 2, 6), (6, 7, 8)),
]


@pytest.mark.parametrize("alg", ["fft", "zncc", "nsqecc"])
@pytest.mark.parametrize("shape,isz,sm,st", CASES)
@pytest.mark.parametrize("mode", ["auto", "fp32", "fp64"])
def test_3d_parity(oracle, alg, shape, isz, sm, st, mode):
    a, b = _vol(shape)
    p = setPIVParameters(corr_alg=alg, interSize=isz, searchMargin=sm, step=st, computeSN=True)
    print(f"case 3d alg={alg} mode={mode} shape={shape} is={isz} sm={sm} st={st}", flush=True)
    VF, SN, PK = PIV(a, b, p, return_peaks=True, mode=mode)
    op = oracle.make_params(alg, 3, isz, sm, st, True, np.float32)
    big = (isz if isinstance(isz, int) else max(isz)) >= 32
    VFo, SNo, aux = oracle.piv(a, b, op, mode=oracle.FFT if big else oracle.DIRECT, nthreads=8,
                               want_aux=True)
    rep = compare(VF, SN, PK, VFo, SNo, aux)
    print(rep, flush=True)
    if alg == "zncc" and mode == "fp32":
        assert rep.peak_mismatch_clear == 0 and rep.vf_maxerr <= 0.01, rep
    else:
        assert rep.ok(), rep


@pytest.mark.parametrize("alg,shape", [
    ("zncc", (256, 16, 128)),    # table pair, dim 1: 257 x 129 columns (one sweep), dim 2 chunked
    ("zncc", (256, 128, 12)),    # table pair, dim 2: 257 x 129 columns (one sweep), dim 1 chunked
    ("nsqecc", (512, 8, 128)),   # single table, dim 1: 513 x 129 columns (one sweep)
])
def test_3d_parity_with_column_sweep_integral_arrays(oracle, alg, shape):
    """Volumes whose integral arrays take the one-sweep column scan (at least
    65536 columns over the tables built together) on one dim and the chunked
    scan on the other."""
    a, b = _vol(shape, density=0.02)
    p = setPIVParameters(corr_alg=alg, interSize=8, searchMargin=2, step=8, computeSN=True)
    print(f"case 3d sweep alg={alg} shape={shape}", flush=True)
    VF, SN, PK = PIV(a, b, p, return_peaks=True, mode="fp64")
    op = oracle.make_params(alg, 3, 8, 2, 8, True, np.float32)
    VFo, SNo, aux = oracle.piv(a, b, op, mode=oracle.DIRECT, nthreads=8, want_aux=True)
    rep = compare(VF, SN, PK, VFo, SNo, aux)
    print(rep, flush=True)
    assert rep.ok(), rep


def test_3d_shear_recovers_displacement(oracle):
    """Linear shear u = 0.01 z + drift on a 64^3 volume: interior vectors follow it."""
    shape = (64, 64, 64)
    n = synth.n_particles(shape, 0.01)
    a, b = synth.particle_pair(shape, n, 1.5, synth.shear3d(), 31, dtype=torch.float32)
    p = setPIVParameters(corr_alg="nsqecc", interSize=16, searchMargin=8, step=16, computeSN=True)
    VF, SN = PIV(a.numpy(), b.numpy(), p)
    # IA (1,1,1) spans z in [16, 32): u ~ 2.3 + 0.01 * 23.5
    assert abs(VF[0, 1, 1, 1] - (2.3 + 0.235)) < 0.15
    assert abs(VF[1, 1, 1, 1] + 1.1) < 0.15
    assert abs(VF[2, 1, 1, 1] - 0.7) < 0.15
