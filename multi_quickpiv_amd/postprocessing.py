"""Vector-field post-processing on the device: the reference's
src/postprocessing/postprocessing.jl filters under the same names, run by the
HIP kernels of libmqpiv (``mqpiv_pp_*``).

Fields use the reference layout: ``VF`` has shape ``(N, nVF...)`` (component
first), per-vector maps have shape ``nVF...``.  Host numpy arrays go in and
come out; the work runs on ``cuda:device``.  ``fun``/``cmp`` arguments that are
Python callables are evaluated on the host (a C ABI cannot call back into the
caller's language); the default magnitude map and the ``>``/``<`` comparisons
run on the device.

Where the reference raises, the intended semantics are used (each is noted on
the function): ``meanReplace`` returns the median, as its body does;
``TH_filtering`` and ``remove_low_SN`` replace as documented; 3-D
``local_magnitude_mask`` is the natural extension.
"""
from __future__ import annotations

import ctypes as C
import math
from decimal import Decimal, getcontext
from typing import Any, Callable, Optional, Tuple, Union

import numpy as np

from . import _lib
from .params import toDims3

__all__ = [
    "magnitudes", "speeds", "velocities", "STD_masking", "TH_masking", "zeroReplace", "medianReplace",
    "meanReplace", "STD_filtering", "TH_filtering", "remove_low_SN", "average", "collectiveness",
    "similarityAveraging", "local_magnitude_mask", "cosd",
]


# ----------------------------------------------------------------- helpers
def cosd(x: float) -> float:
    """cos(x degrees), correctly rounded to double (Taylor series in 50-digit
    decimal arithmetic after exact reduction to [0, 45] degrees)."""
    x = float(x)
    if not math.isfinite(x):
        return math.nan
    getcontext().prec = 60
    xr = Decimal(x) % 360
    if xr < 0:
        xr += 360
    # cos is even and 360-periodic: fold to [0, 180], then use symmetries
    if xr > 180:
        xr = 360 - xr
    sign = 1
    if xr > 90:
        xr = 180 - xr
        sign = -1
    pi = Decimal("3.14159265358979323846264338327950288419716939937510582097494459")
    use_sin = xr > 45
    t = (90 - xr if use_sin else xr) * pi / 180
    term, acc, k = Decimal(1), Decimal(0), 0
    if use_sin:
        term = t
        k = 1
    t2 = t * t
    while True:
        acc += term
        term = -term * t2 / ((k + 1) * (k + 2))
        k += 2
        if abs(term) < Decimal("1e-55"):
            break
    return sign * float(acc)


def _field(VF) -> Tuple[np.ndarray, int]:
    VF = np.asarray(VF)
    if VF.dtype not in (np.float32, np.float64):
        VF = VF.astype(np.float64)
    nd = VF.shape[0] if VF.ndim else 0
    assert 1 < nd < 4 and VF.ndim == nd + 1, "VF must have shape (N, nVF...) with N = 2 or 3"
    return np.asfortranarray(VF), nd


def _desc(nd: int, dtype: np.dtype, shape) -> _lib.MqpivVfDesc:
    d = _lib.MqpivVfDesc()
    d.ndims = nd
    d.dtype = _lib.MQPIV_F32 if np.dtype(dtype) == np.float32 else _lib.MQPIV_F64
    for k in range(3):
        d.size[k] = shape[k] if k < len(shape) else 1
    return d


def _radius(radius, nd: int) -> "C.Array":
    r = toDims3(radius)
    if r is None:
        raise TypeError("radius must be an integer or a tuple of 2 or 3 integers")
    return (C.c_int64 * 3)(*[int(v) for v in r[:nd]], *([0] * (3 - nd)))


_streams: dict = {}


class _Dev:
    """Device staging for one call: uploads, allocations, download and the
    kernels all run on one dedicated (default-priority) stream per device, which
    is passed to the library explicitly -- a NULL stream would select the
    context's own stream, which the copies here are not ordered with."""

    def __init__(self, device: int = 0):
        import torch

        self.torch = torch
        self.dev = torch.device("cuda", device)
        self.ctx = _lib.context(device)
        if device not in _streams:
            _streams[device] = torch.cuda.Stream(self.dev)
        self.s = _streams[device]
        self.stream = self.s.cuda_stream

    def up(self, a: np.ndarray):
        with self.torch.cuda.stream(self.s):
            return self.torch.from_numpy(np.ascontiguousarray(a.ravel(order="F"))).to(self.dev)

    def empty(self, n: int, dtype):
        tdt = {np.dtype(np.float32): self.torch.float32, np.dtype(np.float64): self.torch.float64,
               np.dtype(np.uint8): self.torch.uint8}[np.dtype(dtype)]
        with self.torch.cuda.stream(self.s):
            return self.torch.empty(max(n, 1), dtype=tdt, device=self.dev)

    def down(self, t, shape, n: int) -> np.ndarray:
        self.s.synchronize()
        return t[:n].cpu().numpy().reshape(shape, order="F")

    def check(self, rc: int) -> None:
        self.ctx.check(rc)


def _map_of(VF: np.ndarray, fun: Optional[Callable]) -> Optional[np.ndarray]:
    """The per-vector map a masking step works on: None = magnitudes (device)."""
    if fun is None:
        return None
    m = np.asarray(fun(VF))
    if m.dtype not in (np.float32, np.float64):
        m = m.astype(VF.dtype)
    return np.asfortranarray(m)


# ----------------------------------------------------------------- magnitudes
def magnitudes(VF, typ=None, device: int = 0) -> np.ndarray:
    """magnitudes(VF) (postprocessing.jl:178-194): sqrt(sum(VF[:,i].^2))."""
    VF, nd = _field(VF)
    d = _Dev(device)
    shape = VF.shape[1:]
    n = int(np.prod(shape))
    dv = d.up(VF)
    M = d.empty(n, VF.dtype)
    d.check(d.ctx.lib.mqpiv_pp_magnitudes(d.ctx.handle, C.byref(_desc(nd, VF.dtype, shape)), dv.data_ptr(),
                                          M.data_ptr(), d.stream))
    out = d.down(M, shape, n)
    return out.astype(typ) if typ is not None else out


speeds = magnitudes
velocities = magnitudes


def _device_map(d: _Dev, VF: np.ndarray, nd: int, fun):
    """Uploads the map (or computes magnitudes on the device); returns (tensor, dtype)."""
    shape = VF.shape[1:]
    n = int(np.prod(shape))
    m = _map_of(VF, fun)
    if m is not None:
        assert m.shape == shape, "fun(VF) must return one value per vector"
        return d.up(m), m.dtype
    dv = d.up(VF)
    M = d.empty(n, VF.dtype)
    d.check(d.ctx.lib.mqpiv_pp_magnitudes(d.ctx.handle, C.byref(_desc(nd, VF.dtype, shape)), dv.data_ptr(),
                                          M.data_ptr(), d.stream))
    return M, VF.dtype


def _std_mask_dev(d: _Dev, VF, nd, n_std, fun):
    shape = VF.shape[1:]
    n = int(np.prod(shape))
    m, mdt = _device_map(d, VF, nd, fun)
    mask = d.empty(n, np.uint8)
    d.check(d.ctx.lib.mqpiv_pp_std_mask(d.ctx.handle, C.byref(_desc(nd, mdt, shape)), m.data_ptr(),
                                        float(n_std), mask.data_ptr(), d.stream))
    return mask


def _cmp_code(cmp) -> Optional[int]:
    if cmp is None or (isinstance(cmp, str) and cmp == ">"):
        return _lib.MQPIV_CMP_GT
    if isinstance(cmp, str) and cmp == "<":
        return _lib.MQPIV_CMP_LT
    if isinstance(cmp, str):
        raise ValueError(f"cmp must be '>', '<' or a callable, not {cmp!r}")
    return None  # a callable: host-evaluated


def _th_mask_dev(d: _Dev, VF, nd, th, fun, cmp):
    shape = VF.shape[1:]
    n = int(np.prod(shape))
    code = _cmp_code(cmp)
    if code is None:
        mp = _map_of(VF, fun)
        if mp is None:
            mp = magnitudes(VF, device=d.dev.index or 0)
        host = np.vectorize(lambda x: bool(cmp(x, th)), otypes=[bool])(mp).astype(np.uint8)
        return d.up(np.asfortranarray(host))
    m, mdt = _device_map(d, VF, nd, fun)
    mask = d.empty(n, np.uint8)
    d.check(d.ctx.lib.mqpiv_pp_th_mask(d.ctx.handle, C.byref(_desc(nd, mdt, shape)), m.data_ptr(), float(th),
                                       code, mask.data_ptr(), d.stream))
    return mask


# ----------------------------------------------------------------- masking
def STD_masking(n, VF, fun: Optional[Callable] = None, device: int = 0) -> np.ndarray:
    """STD_masking(n, VF; fun=magnitudes) (postprocessing.jl:598-610): Bool map of
    |map - mean| > n*std (corrected std).  Pass ``fun=lambda x: x`` semantics by
    giving a per-vector array as ``fun(VF)``."""
    VF, nd = _field(VF)
    d = _Dev(device)
    shape = VF.shape[1:]
    mask = _std_mask_dev(d, VF, nd, n, fun)
    return d.down(mask, shape, int(np.prod(shape))).astype(bool)


def TH_masking(th, VF, fun: Optional[Callable] = None, cmp: Union[None, str, Callable] = None,
               device: int = 0) -> np.ndarray:
    """TH_masking(th, VF; fun=magnitudes, cmp=(x,th)->x>th) (postprocessing.jl:640-663).
    ``cmp`` may be ">" (default), "<" or a callable (host-evaluated)."""
    VF, nd = _field(VF)
    d = _Dev(device)
    shape = VF.shape[1:]
    mask = _th_mask_dev(d, VF, nd, th, fun, cmp)
    return d.down(mask, shape, int(np.prod(shape))).astype(bool)


# ----------------------------------------------------------------- replacement
def zeroReplace(coords, radius, VF):  # noqa: N802
    """zeroReplace (postprocessing.jl:539-545): the zero vector."""
    return np.zeros(np.shape(VF)[0], dtype=np.asarray(VF).dtype)


def medianReplace(coords, radius, VF):  # noqa: N802
    """medianReplace (postprocessing.jl:547-553): per-component median of the
    clipped box around ``coords`` (1-based)."""
    VF = np.asarray(VF)
    nd = VF.shape[0]
    r = toDims3(radius)[:nd]
    sl = tuple(slice(max(1, c - rr) - 1, min(s, c + rr)) for c, rr, s in zip(coords, r, VF.shape[1:]))
    return np.array([np.median(VF[(k,) + sl]) for k in range(nd)], dtype=VF.dtype)


def meanReplace(coords, radius, VF):  # noqa: N802
    """meanReplace (postprocessing.jl:555-561): its body takes the MEDIAN.  (Its
    VF method splats VF into a tuple that localOp has no method for, so the
    reference throws; the body's semantics are used.)"""
    return medianReplace(coords, radius, VF)


def _replace_code(f) -> Optional[int]:
    if f is zeroReplace:
        return _lib.MQPIV_REPLACE_ZERO
    if f is medianReplace or f is meanReplace:
        return _lib.MQPIV_REPLACE_MEDIAN
    return None


def _filter(VF: np.ndarray, nd: int, d: _Dev, mask_t, replaceFun, radius) -> np.ndarray:
    shape = VF.shape[1:]
    n = int(np.prod(shape))
    how = _replace_code(replaceFun)
    if how is None:  # a caller-supplied replacement: host loop over flagged vectors
        mask = d.down(mask_t, shape, n).astype(bool)
        out = VF.copy(order="F")
        for idx in zip(*np.nonzero(mask)):
            out[(slice(None),) + idx] = replaceFun(tuple(i + 1 for i in idx), radius, VF)
        return out
    dv = d.up(VF)
    out = d.empty(n * nd, VF.dtype)
    d.check(d.ctx.lib.mqpiv_pp_replace(d.ctx.handle, C.byref(_desc(nd, VF.dtype, shape)), dv.data_ptr(),
                                       mask_t.data_ptr(), how, _radius(radius, nd), out.data_ptr(), d.stream))
    return d.down(out, VF.shape, n * nd)


def STD_filtering(n, VF, fun: Optional[Callable] = None, replaceFun: Callable = meanReplace,  # noqa: N802
                  radius=2, device: int = 0) -> np.ndarray:
    """STD_filtering(n, VF; fun, replaceFun=meanReplace, radius=2)
    (postprocessing.jl:618-634): STD_masking, then every flagged vector replaced
    (computed from the unfiltered field)."""
    VF, nd = _field(VF)
    d = _Dev(device)
    return _filter(VF, nd, d, _std_mask_dev(d, VF, nd, n, fun), replaceFun, radius)


def TH_filtering(th, VF, fun: Optional[Callable] = None, cmp: Union[None, str, Callable] = None,  # noqa: N802
                 replaceFun: Callable = meanReplace, radius=2, device: int = 0) -> np.ndarray:
    """TH_filtering(th, VF; fun, cmp, replaceFun=meanReplace, radius=2)
    (postprocessing.jl:672-703).  The reference calls an undefined
    ``replacefun`` once a vector is flagged; here ``replaceFun`` is applied."""
    VF, nd = _field(VF)
    d = _Dev(device)
    return _filter(VF, nd, d, _th_mask_dev(d, VF, nd, th, fun, cmp), replaceFun, radius)


def remove_low_SN(SN, threshold, VF, replace: Callable = zeroReplace, radius=1,  # noqa: N802
                  device: int = 0) -> np.ndarray:
    """remove_low_SN(SN, threshold, VF; replace=zeroReplace, radius=1)
    (postprocessing.jl:712-732): vectors with SN < threshold replaced.  (The
    reference forwards an unsupported ``replace`` keyword to TH_filtering.)"""
    VF, nd = _field(VF)
    SN = np.asfortranarray(np.asarray(SN, dtype=VF.dtype))
    assert SN.shape == VF.shape[1:], "SN must have one value per vector"
    d = _Dev(device)
    return _filter(VF, nd, d, _th_mask_dev(d, VF, nd, threshold, lambda _x: SN, "<"), replace, radius)


# ----------------------------------------------------------------- smoothing
def average(radius, VF, device: int = 0) -> np.ndarray:
    """average(radius, VF) (postprocessing.jl:359-381): per-component box mean."""
    VF, nd = _field(VF)
    d = _Dev(device)
    shape = VF.shape[1:]
    n = int(np.prod(shape)) * nd
    dv = d.up(VF)
    out = d.empty(n, VF.dtype)
    d.check(d.ctx.lib.mqpiv_pp_average(d.ctx.handle, C.byref(_desc(nd, VF.dtype, shape)), dv.data_ptr(),
                                       _radius(radius, nd), out.data_ptr(), d.stream))
    return d.down(out, VF.shape, n)


def collectiveness(radius, VF, max_angle=40, device: int = 0) -> np.ndarray:
    """collectiveness(radius, VF; max_angle=40) (postprocessing.jl:294-331):
    Float32 fraction of the clipped box whose direction is within max_angle."""
    VF, nd = _field(VF)
    d = _Dev(device)
    shape = VF.shape[1:]
    n = int(np.prod(shape))
    dv = d.up(VF)
    out = d.empty(n, np.float32)
    d.check(d.ctx.lib.mqpiv_pp_collectiveness(d.ctx.handle, C.byref(_desc(nd, VF.dtype, shape)), dv.data_ptr(),
                                              _radius(radius, nd), cosd(max_angle), out.data_ptr(), d.stream))
    return d.down(out, shape, n)


def similarityAveraging(radius, VF, max_angle=20, normalize: bool = True,  # noqa: N802
                        scale_by_similarity: bool = True, similarity_exponent: int = 2,
                        device: int = 0) -> np.ndarray:
    """similarityAveraging(radius, VF; max_angle=20, normalize=true,
    scale_by_similarity=true, similarity_exponent=2) (postprocessing.jl:478-523)."""
    VF, nd = _field(VF)
    if int(similarity_exponent) != similarity_exponent or similarity_exponent < 0:
        raise NotImplementedError("similarity_exponent must be a non-negative integer")
    d = _Dev(device)
    shape = VF.shape[1:]
    n = int(np.prod(shape)) * nd
    dv = d.up(VF)
    out = d.empty(n, VF.dtype)
    d.check(d.ctx.lib.mqpiv_pp_similarity_averaging(
        d.ctx.handle, C.byref(_desc(nd, VF.dtype, shape)), dv.data_ptr(), _radius(radius, nd), cosd(max_angle),
        int(bool(normalize)), int(bool(scale_by_similarity)), int(similarity_exponent), out.data_ptr(),
        d.stream))
    return d.down(out, VF.shape, n)


def local_magnitude_mask(ratio, VF, radius=4, device: int = 0) -> Tuple[np.ndarray, np.ndarray]:
    """local_magnitude_mask(ratio, VF; radius=4) (postprocessing.jl:202-218) ->
    (VF_c, removed).  (The reference's VF_c[c,:,:] indexing throws for 3-D
    fields; the 3-D form is the natural extension.)"""
    VF, nd = _field(VF)
    d = _Dev(device)
    shape = VF.shape[1:]
    nv = int(np.prod(shape))
    dv = d.up(VF)
    out = d.empty(nv * nd, VF.dtype)
    removed = d.empty(nv, np.uint8)
    d.check(d.ctx.lib.mqpiv_pp_local_magnitude_mask(
        d.ctx.handle, C.byref(_desc(nd, VF.dtype, shape)), dv.data_ptr(), float(ratio), _radius(radius, nd),
        out.data_ptr(), removed.data_ptr(), d.stream))
    return d.down(out, VF.shape, nv * nd), d.down(removed, shape, nv).astype(bool)
