"""Python access to the Float64 CPU oracle (oracle/mqpiv_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as the
checker, and by bench.py for the timed CPU baseline.  The product package never
imports it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path
from typing import Optional, Sequence, Tuple

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liborc.so"

FFTCC, ZNCC, NSQECC, MASK_NSQECC = 0, 1, 2, 3
DIRECT, FFT = 0, 1
ALG_CODES = {"fft": FFTCC, "fftcc": FFTCC, "zncc": ZNCC, "nsqecc": NSQECC, "mask_nsqecc": MASK_NSQECC}


class OrcParams(C.Structure):
    _fields_ = [
        ("alg", C.c_int32),
        ("ndims", C.c_int32),
        ("is_", C.c_int64 * 3),
        ("sm", C.c_int64 * 3),
        ("st", C.c_int64 * 3),
        ("compute_sn", C.c_int32),
        ("in_dtype", C.c_int32),
        ("threshold", C.c_double),
        ("mask_threshold", C.c_double),
    ]


_lib = None


def load() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
        lib = C.CDLL(str(LIB_PATH))
        vp = C.c_void_p
        lib.orc_good_size_real.restype = C.c_int64
        lib.orc_good_size_real.argtypes = [C.c_int64]
        lib.orc_vf_size.restype = C.c_int64
        lib.orc_vf_size.argtypes = [C.POINTER(OrcParams), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        lib.orc_piv.restype = C.c_int
        lib.orc_piv.argtypes = [C.POINTER(OrcParams), C.POINTER(C.c_int64), vp, vp, vp, vp, C.c_int,
                                C.c_int, C.c_int64, C.c_int64, vp, vp, vp]
        lib.orc_skip_flags.restype = C.c_int
        lib.orc_skip_flags.argtypes = [C.POINTER(OrcParams), C.POINTER(C.c_int64), vp, vp, vp]
        lib.orc_piv_multipass.restype = C.c_int
        lib.orc_piv_multipass.argtypes = [C.POINTER(OrcParams), C.c_int, C.POINTER(C.c_int64), vp, vp, vp,
                                          C.c_int, C.c_int, vp, vp]
        lib.orc_coords_shifted.restype = None
        lib.orc_coords_shifted.argtypes = [C.POINTER(OrcParams), C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                           C.c_int64, C.POINTER(C.c_int64), vp]
        lib.orc_shift_field.restype = None
        lib.orc_shift_field.argtypes = [C.c_int, vp] + [C.POINTER(C.c_int64)] * 6 + [vp]
        lib.orc_ia_similarity.restype = C.c_int
        lib.orc_ia_similarity.argtypes = [C.POINTER(OrcParams), C.POINTER(C.c_int64), vp, vp, vp,
                                          C.c_int64, C.c_int, vp, vp]
        lib.orc_ia_correlation.restype = C.c_int
        lib.orc_ia_correlation.argtypes = [C.POINTER(OrcParams), C.POINTER(C.c_int64), vp, vp,
                                           C.c_int64, C.c_int, vp]
        lib.orc_coords_of.restype = None
        lib.orc_coords_of.argtypes = [C.POINTER(OrcParams), C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                      C.c_int64, vp]
        lib.orc_peak.restype = None
        lib.orc_peak.argtypes = [vp, C.POINTER(C.c_int64), C.c_int, C.POINTER(C.c_int64),
                                 C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                 C.POINTER(C.c_double)]
        lib.orc_subpixel.restype = None
        lib.orc_subpixel.argtypes = [vp, C.POINTER(C.c_int64), C.c_int, C.POINTER(C.c_int64),
                                     C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                     C.c_double, C.POINTER(C.c_double)]
        lib.orc_set_plant.restype = None
        lib.orc_set_plant.argtypes = [C.c_int]
        lib.orc_sn.restype = C.c_double
        lib.orc_sn.argtypes = [vp, C.POINTER(C.c_int64), C.c_int, C.POINTER(C.c_int64),
                               C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        _lib = lib
    return _lib


def _i3(v: Sequence[int], fill: int) -> "C.Array":
    v = list(v) + [fill] * (3 - len(v))
    return (C.c_int64 * 3)(*v)


def make_params(alg, ndims: int, inter_size, search_margin, step, compute_sn: bool = True,
                in_dtype=np.float64, threshold: float = -1.0, mask_threshold: float = -1.0) -> OrcParams:
    p = OrcParams()
    p.alg = ALG_CODES[alg.lower()] if isinstance(alg, str) else int(alg)
    p.ndims = ndims
    tup = lambda x: (x,) * ndims if isinstance(x, (int, np.integer)) else tuple(x)
    p.is_ = _i3(tup(inter_size), 1)
    p.sm = _i3(tup(search_margin), 0)
    p.st = _i3(tup(step), 1)
    p.compute_sn = 1 if compute_sn else 0
    p.in_dtype = 0 if np.dtype(in_dtype) == np.float32 else 1
    p.threshold = threshold
    p.mask_threshold = mask_threshold
    return p


# Planted mistakes (orc_set_plant in mqpiv_oracle.c): used by the pinning tests
# to show that their checks detect each one.
PLANT_NONE, PLANT_MASK_C1_UNMASKED, PLANT_MASK_C2_BOXSUM, PLANT_MASK_NO_W, PLANT_ZNCC_RAW_F2 = range(5)


def set_plant(k: int) -> None:
    load().orc_set_plant(int(k))


def good_size_real(n: int) -> int:
    return int(load().orc_good_size_real(n))


def vf_size(p: OrcParams, shape: Sequence[int]) -> Tuple[int, ...]:
    vf = (C.c_int64 * 3)()
    load().orc_vf_size(C.byref(p), _i3(shape, 1), vf)
    return tuple(vf[d] for d in range(p.ndims))


def _prep(img: Optional[np.ndarray], dtype) -> Optional[np.ndarray]:
    if img is None:
        return None
    return np.asfortranarray(img, dtype=dtype)


def piv(img1, img2, p: OrcParams, mask=None, skip_mask=None, mode: int = DIRECT, nthreads: int = 1,
        ia_begin: int = 0, ia_end: int = -1, want_aux: bool = False):
    """PIV_CPU restated in Float64.  Returns (VF, SN, aux) with VF of shape
    (N, nVF...), SN of shape nVF (or None), aux of shape (6, nVF...) holding the
    1-based integer peak (r, c, z), the peak value and the runner-up value in the
    peak window (or None)."""
    lib = load()
    dt = np.float32 if p.in_dtype == 0 else np.float64
    a1, a2, am = _prep(img1, dt), _prep(img2, dt), _prep(mask, dt)
    shape = a1.shape
    vf = vf_size(p, shape)
    n = int(np.prod(vf))
    VF = np.zeros((p.ndims, *vf), dtype=np.float64, order="F")
    SN = np.zeros(vf, dtype=np.float64, order="F") if p.compute_sn else None
    aux = np.zeros((6, *vf), dtype=np.float64, order="F") if want_aux else None
    sk = np.ascontiguousarray(skip_mask, dtype=np.uint8) if skip_mask is not None else None
    if n:
        rc = lib.orc_piv(C.byref(p), _i3(shape, 1), a1.ctypes.data, a2.ctypes.data,
                         am.ctypes.data if am is not None else None,
                         sk.ctypes.data if sk is not None else None, mode, nthreads, ia_begin, ia_end,
                         VF.ctypes.data, SN.ctypes.data if SN is not None else None,
                         aux.ctypes.data if aux is not None else None)
        if rc != 0:
            raise ValueError(f"orc_piv failed ({rc})")
    return VF, SN, aux


def skip_flags(img1, p: OrcParams, mask=None) -> np.ndarray:
    """Skip flags of skip_inter_region for every IA (uint8, 1 = skipped, IA order)
    read from image 1 and, for the masked measure, ``mask`` (mask_VF in the
    two-mask driver)."""
    lib = load()
    dt = np.float32 if p.in_dtype == 0 else np.float64
    a1, am = _prep(img1, dt), _prep(mask, dt)
    n = int(np.prod(vf_size(p, a1.shape)))
    out = np.zeros(max(n, 1), dtype=np.uint8)
    rc = lib.orc_skip_flags(C.byref(p), _i3(a1.shape, 1), a1.ctypes.data,
                            am.ctypes.data if am is not None else None, out.ctypes.data)
    if rc != 0:
        raise ValueError(f"orc_skip_flags failed ({rc})")
    return out[:n]


def piv_multipass(img1, img2, p: OrcParams, multipass: int, mask=None, mode: int = DIRECT,
                  nthreads: int = 1):
    """Multipass PIV (orc_piv_multipass): (VF, SN) of the last pass."""
    lib = load()
    dt = np.float32 if p.in_dtype == 0 else np.float64
    a1, a2, am = _prep(img1, dt), _prep(img2, dt), _prep(mask, dt)
    shape = a1.shape
    vf = vf_size(p, shape)
    VF = np.zeros((p.ndims, *vf), dtype=np.float64, order="F")
    SN = np.zeros(vf, dtype=np.float64, order="F") if p.compute_sn else None
    rc = lib.orc_piv_multipass(C.byref(p), multipass, _i3(shape, 1), a1.ctypes.data, a2.ctypes.data,
                               am.ctypes.data if am is not None else None, mode, nthreads, VF.ctypes.data,
                               SN.ctypes.data if SN is not None else None)
    if rc != 0:
        raise ValueError(f"orc_piv_multipass failed ({rc})")
    return VF, SN


def coords_shifted(p: OrcParams, shape: Sequence[int], vf_idx: int, d: Sequence[int]):
    """(coords, clamped d): the 18 coordinates (TLF, BRB, SR_TLF, SR_BRB,
    SR_TLF_off, SR_BRB_off; 3 each) of IA vf_idx (1-based) whose search region
    is shifted by d, and the shift after clamping."""
    lib = load()
    vf = (C.c_int64 * 3)(*(list(vf_size(p, shape)) + [1] * (3 - p.ndims)))
    dd = _i3(d, 0)
    buf = np.zeros(18, dtype=np.int64)
    lib.orc_coords_shifted(C.byref(p), _i3(shape, 1), vf, vf_idx, dd, buf.ctypes.data)
    return buf.reshape(6, 3), tuple(dd[k] for k in range(p.ndims))


def shift_field(vf_coarse: np.ndarray, is_c, st_c, vf_fine_shape, is_f, st_f) -> np.ndarray:
    """Search-region shifts (shape (3, nVF_fine...), int32) that a pass with IA
    size is_f and step st_f takes from the coarse field vf_coarse (N, nVF_c...)."""
    nd = vf_coarse.shape[0]
    vc = np.asfortranarray(vf_coarse, dtype=np.float64)
    out = np.zeros((3, *vf_fine_shape), dtype=np.int32, order="F")
    load().orc_shift_field(nd, vc.ctypes.data, _i3(vf_coarse.shape[1:], 1), _i3(is_c, 1), _i3(st_c, 1),
                           _i3(vf_fine_shape, 1), _i3(is_f, 1), _i3(st_f, 1), out.ctypes.data)
    return out


def ia_similarity(img1, img2, p: OrcParams, vf_idx: int, mask=None, mode: int = DIRECT):
    """Similarity window of the IA with 1-based linear index vf_idx, shape
    (2SM+1,)*N in column-major order, and its 18 coordinates
    (TLF, BRB, SR_TLF, SR_BRB, SR_TLF_off, SR_BRB_off; 3 each, 1-based)."""
    lib = load()
    dt = np.float32 if p.in_dtype == 0 else np.float64
    a1, a2, am = _prep(img1, dt), _prep(img2, dt), _prep(mask, dt)
    ns = tuple(2 * p.sm[d] + 1 for d in range(p.ndims))
    out = np.zeros(ns, dtype=np.float64, order="F")
    co = np.zeros(18, dtype=np.int64)
    lib.orc_ia_similarity(C.byref(p), _i3(a1.shape, 1), a1.ctypes.data, a2.ctypes.data,
                          am.ctypes.data if am is not None else None, vf_idx, mode, out.ctypes.data,
                          co.ctypes.data)
    return out, co.reshape(6, 3)


def ia_correlation(img1, img2, p: OrcParams, vf_idx: int, mode: int = DIRECT):
    lib = load()
    dt = np.float32 if p.in_dtype == 0 else np.float64
    a1, a2 = _prep(img1, dt), _prep(img2, dt)
    ns = tuple(2 * p.sm[d] + 1 for d in range(p.ndims))
    out = np.zeros(ns, dtype=np.float64, order="F")
    lib.orc_ia_correlation(C.byref(p), _i3(a1.shape, 1), a1.ctypes.data, a2.ctypes.data, vf_idx, mode,
                           out.ctypes.data)
    return out


def coords(p: OrcParams, shape: Sequence[int], vf_idx: int) -> np.ndarray:
    lib = load()
    vf = (C.c_int64 * 3)()
    lib.orc_vf_size(C.byref(p), _i3(shape, 1), vf)
    co = np.zeros(18, dtype=np.int64)
    lib.orc_coords_of(C.byref(p), _i3(shape, 1), vf, vf_idx, co.ctypes.data)
    return co.reshape(6, 3)


def peak_subpixel_sn(S: np.ndarray, sm: Sequence[int], off_lo: Sequence[int], off_hi: Sequence[int]):
    """first_fullovp_peak + gaussian refinement + compute_SN on a given window."""
    lib = load()
    S = np.asfortranarray(S, dtype=np.float64)
    nd = S.ndim
    ns = _i3(S.shape, 1)
    sm3, lo3, hi3 = _i3(sm, 0), _i3(off_lo, 0), _i3(off_hi, 0)
    pk = (C.c_int64 * 3)()
    val = C.c_double()
    lib.orc_peak(S.ctypes.data, ns, nd, sm3, lo3, hi3, pk, C.byref(val))
    disp = (C.c_double * 3)()
    lib.orc_subpixel(S.ctypes.data, ns, nd, sm3, lo3, hi3, pk, val.value, disp)
    sn = lib.orc_sn(S.ctypes.data, ns, nd, sm3, lo3, hi3)
    return tuple(pk[d] for d in range(nd)), val.value, tuple(disp[d] for d in range(nd)), sn


# ---------------------------------------------------------------- post-processing
# Restatement of src/postprocessing/postprocessing.jl (oracle/mqpiv_oracle_pp.c).
def _sfx(a: np.ndarray) -> str:
    return "_f32" if a.dtype == np.float32 else "_f64"


def _vf_in(VF) -> Tuple[np.ndarray, int, "C.Array"]:
    VF = np.asarray(VF)
    if VF.dtype not in (np.float32, np.float64):
        VF = VF.astype(np.float64)
    VF = np.asfortranarray(VF)
    nd = VF.shape[0]
    assert VF.ndim == nd + 1 and nd in (2, 3), "VF must be (N, nVF...) with N = 2 or 3"
    return VF, nd, _i3(VF.shape[1:], 1)


def _r3(radius, nd: int) -> "C.Array":
    r = (radius,) * nd if isinstance(radius, (int, np.integer)) else tuple(radius)[:nd]
    return _i3(r, 0)


def _fn(name: str, a: np.ndarray):
    return getattr(load(), f"orc_pp_{name}{_sfx(a)}")


def pp_magnitudes(VF) -> np.ndarray:
    VF, nd, sz = _vf_in(VF)
    M = np.zeros(VF.shape[1:], dtype=VF.dtype, order="F")
    f = _fn("magnitudes", VF)
    f.restype = None
    f(C.c_int(nd), sz, C.c_void_p(VF.ctypes.data), C.c_void_p(M.ctypes.data))
    return M


def pp_th_mask(x, th: float, cmp: int = 0) -> np.ndarray:
    x = np.asfortranarray(x)
    mask = np.zeros(x.shape, dtype=np.uint8, order="F")
    f = _fn("th_mask", x)
    f.restype = None
    f(C.c_int64(x.size), C.c_void_p(x.ctypes.data), C.c_double(th), C.c_int(cmp), C.c_void_p(mask.ctypes.data))
    return mask


def pp_std_mask(x, nsd: float):
    """(mask, mean, std) of STD_masking on a per-vector map."""
    x = np.asfortranarray(x)
    mask = np.zeros(x.shape, dtype=np.uint8, order="F")
    stats = (C.c_double * 2)()
    f = _fn("std_mask", x)
    f.restype = None
    f(C.c_int64(x.size), C.c_void_p(x.ctypes.data), C.c_double(nsd), C.c_void_p(mask.ctypes.data), stats)
    return mask, stats[0], stats[1]


def pp_replace(VF, mask, how: int, radius) -> np.ndarray:
    VF, nd, sz = _vf_in(VF)
    mask = np.asfortranarray(mask, dtype=np.uint8)
    out = np.zeros_like(VF, order="F")
    f = _fn("replace", VF)
    f.restype = None
    f(C.c_int(nd), sz, C.c_void_p(VF.ctypes.data), C.c_void_p(mask.ctypes.data), C.c_int(how),
      _r3(radius, nd), C.c_void_p(out.ctypes.data))
    return out


def pp_average(VF, radius) -> np.ndarray:
    VF, nd, sz = _vf_in(VF)
    out = np.zeros_like(VF, order="F")
    f = _fn("average", VF)
    f.restype = None
    f(C.c_int(nd), sz, C.c_void_p(VF.ctypes.data), _r3(radius, nd), C.c_void_p(out.ctypes.data))
    return out


def pp_collectiveness(VF, radius, min_dot: float) -> np.ndarray:
    VF, nd, sz = _vf_in(VF)
    out = np.zeros(VF.shape[1:], dtype=np.float32, order="F")
    f = _fn("collectiveness", VF)
    f.restype = None
    f(C.c_int(nd), sz, C.c_void_p(VF.ctypes.data), _r3(radius, nd), C.c_double(min_dot),
      C.c_void_p(out.ctypes.data))
    return out


def pp_simavg(VF, radius, min_dot: float, normalize: bool = True, scale: bool = True, expo: int = 2):
    VF, nd, sz = _vf_in(VF)
    out = np.zeros_like(VF, order="F")
    f = _fn("simavg", VF)
    f.restype = None
    f(C.c_int(nd), sz, C.c_void_p(VF.ctypes.data), _r3(radius, nd), C.c_double(min_dot),
      C.c_int(int(normalize)), C.c_int(int(scale)), C.c_int(expo), C.c_void_p(out.ctypes.data))
    return out


def pp_lmm(VF, ratio: float, radius):
    VF, nd, sz = _vf_in(VF)
    out = np.zeros_like(VF, order="F")
    removed = np.zeros(VF.shape[1:], dtype=np.uint8, order="F")
    f = _fn("lmm", VF)
    f.restype = None
    f(C.c_int(nd), sz, C.c_void_p(VF.ctypes.data), C.c_double(ratio), _r3(radius, nd),
      C.c_void_p(out.ctypes.data), C.c_void_p(removed.ctypes.data))
    return out, removed
