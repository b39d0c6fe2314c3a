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


def _ellipsoid_mask(shape, n, seed):
    """1 outside n seeded random ellipsoids (about 20-30% of the volume), 0 inside."""
    r = np.random.default_rng(seed)
    g = np.meshgrid(*[np.arange(s, dtype=np.float64) for s in shape], indexing="ij")
    m = np.ones(shape, dtype=np.float32)
    for _ in range(n):
        c = [r.uniform(0, s) for s in shape]
        ax = [r.uniform(0.1, 0.25) * s for s in shape]
        inside = sum(((g[d] - c[d]) / ax[d]) ** 2 for d in range(3)) <= 1.0
        m[inside] = 0.0
    return m


MASK_CASES = [
    ((40, 36, 32), 8, 4, 8),
    ((48, 44, 40), (12, 8, 8), (4, 3, 2), (6, 7, 8)),
    ((64, 64, 48), 16, 8, 12),
]


@pytest.mark.parametrize("shape,isz,sm,st", MASK_CASES)
@pytest.mark.parametrize("mode", ["auto", "fp64"])
def test_3d_masked_nsqecc_parity(oracle, shape, isz, sm, st, mode):
    """Masked NSQECC on volumes (mask_nsqecc.jl:126-160 handles N = 3): the two
    correlations (F m) x G and m x G^2 through the 3-D pipeline, FP64 in AUTO."""
    a, b = _vol(shape, seed=41)
    m = _ellipsoid_mask(shape, 5, 42)
    p = setPIVParameters(corr_alg="nsqecc", interSize=isz, searchMargin=sm, step=st, computeSN=True)
    print(f"case 3d masked mode={mode} shape={shape} is={isz} sm={sm} st={st}", flush=True)
    VF, SN, PK = PIV(a, b, m, p, return_peaks=True, mode=mode)
    op = oracle.make_params("mask_nsqecc", 3, isz, sm, st, True, np.float32)
    VFo, SNo, aux = oracle.piv(a, b, op, mask=m, mode=oracle.DIRECT, nthreads=8, want_aux=True)
    rep = compare(VF, SN, PK, VFo, SNo, aux)
    print(rep, flush=True)
    assert rep.ok(), rep


def test_3d_masked_thresholds_and_double_mask(oracle):
    """3-D masked PIV with a threshold and a mask threshold (IAs skipped on image 1
    or on the mask), single-mask and double-mask forms (piv_ops.jl:148-157,
    multi_quickPIV.jl:155-213): the double-mask skip flags come from the device
    kernel over mask_VF, the correlation uses mask_IA."""
    shape = (40, 36, 32)
    a, b = _vol(shape, seed=43)
    m_ia = _ellipsoid_mask(shape, 4, 44)
    m_vf = _ellipsoid_mask(shape, 6, 45)
    thr = float(np.quantile(a, 0.995))
    p = setPIVParameters(corr_alg="nsqecc", interSize=8, searchMargin=4, step=8, computeSN=True,
                         threshold=thr, mask_threshold=0.5)
    op = oracle.make_params("mask_nsqecc", 3, 8, 4, 8, True, np.float32, thr, 0.5)
    VF, SN, PK = PIV(a, b, m_ia, p, return_peaks=True)
    VFo, SNo, aux = oracle.piv(a, b, op, mask=m_ia, mode=oracle.DIRECT, nthreads=8, want_aux=True)
    rep = compare(VF, SN, PK, VFo, SNo, aux)
    print("single", rep, flush=True)
    skipped = (VFo == 0).all(axis=0) & (SNo == 0)
    assert skipped.any() and (~skipped).any()
    assert rep.ok(), rep
    VF2, SN2, PK2 = PIV(a, b, m_vf, m_ia, p, return_peaks=True)
    flags = oracle.skip_flags(a, op, mask=m_vf)
    VFo2, SNo2, aux2 = oracle.piv(a, b, op, mask=m_ia, skip_mask=flags, mode=oracle.DIRECT, nthreads=8,
                                  want_aux=True)
    rep2 = compare(VF2, SN2, PK2, VFo2, SNo2, aux2)
    print("double", rep2, flush=True)
    assert rep2.ok(), rep2
