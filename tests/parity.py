"""Shared helpers for comparing the HIP path against the Float64 oracle.

Parity bar (stated once, used by every GPU parity test):
  * integer peaks: identical wherever the oracle's top peak beats the runner-up
    in the peak window by more than NEAR_TIE (relative); others are counted;
  * displacement: |VF_gpu - VF_oracle| <= VF_TOL px where the peaks agree;
  * SN: relative difference <= SN_RTOL (NaN matches NaN).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

NEAR_TIE = 1e-6
VF_TOL = 0.01
SN_RTOL = 1e-4


@dataclass
class Report:
    n: int
    n_clear: int
    peak_mismatch_clear: int
    peak_mismatch_near: int
    vf_maxerr: float
    sn_maxrel: float
    n_nopeak: int

    def ok(self) -> bool:
        return (self.peak_mismatch_clear == 0 and self.vf_maxerr <= VF_TOL
                and self.sn_maxrel <= SN_RTOL)


def compare(VF, SN, PK, VFo, SNo, aux) -> Report:
    N = VF.shape[0]
    n = int(np.prod(VF.shape[1:]))
    pk_o = aux[:N].reshape(N, n).T.astype(np.int64)
    pk_g = PK.reshape(N, n).T.astype(np.int64)
    top = aux[3].reshape(n)
    run = aux[4].reshape(n)
    has = np.all(pk_o > 0, axis=1)
    with np.errstate(invalid="ignore", divide="ignore"):
        margin = np.where(has, (top - run) / np.abs(top), np.inf)
    clear = margin > NEAR_TIE
    same = np.all(pk_o == pk_g, axis=1)
    vf_g = VF.reshape(N, n).T.astype(np.float64)
    vf_o = VFo.reshape(N, n).T
    err = np.abs(vf_g - vf_o)
    err = np.where(np.isnan(vf_g) & np.isnan(vf_o), 0.0, err)
    vf_max = float(np.max(err[same])) if same.any() else 0.0
    sn_max = 0.0
    if SN is not None:
        a = SN.reshape(n).astype(np.float64)
        b = SNo.reshape(n)
        both_nan = np.isnan(a) & np.isnan(b)
        with np.errstate(invalid="ignore", divide="ignore"):
            rel = np.abs(a - b) / np.maximum(np.abs(b), 1e-300)
        rel = np.where(both_nan, 0.0, rel)
        rel = np.where(np.isnan(rel), np.inf, rel)
        sn_max = float(np.max(rel)) if n else 0.0
    return Report(
        n=n,
        n_clear=int(clear.sum()),
        peak_mismatch_clear=int((~same & clear).sum()),
        peak_mismatch_near=int((~same & ~clear).sum()),
        vf_maxerr=vf_max,
        sn_maxrel=sn_max,
        n_nopeak=int((~has).sum()),
    )
