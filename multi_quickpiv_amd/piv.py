"""``PIV(img1, img2[, mask[, mask_IA]], pivparams; precision)`` on the MI355X path.

Mirrors src/multi_quickPIV.jl:17-29 (the three ``PIV`` methods) with the
outputs of src/utils/piv_ops.jl:25-31:

* ``VF`` has shape ``(N, nVF_1, ..., nVF_N)``, component first (U = VF[0], ...),
* ``SN`` has shape ``nVF`` or is ``None`` when ``computeSN`` is false.

Arrays use the reference's dimension order: axis 0 is Julia's dim 1 (rows, U).
They are handed to the C ABI in column-major order (dim 1 fastest), exactly the
memory of a Julia array, so VF/SN come back in the reference's element order.
"""
from __future__ import annotations

import copy

import ctypes as C
from typing import Any, Optional, Tuple

import numpy as np

from . import _lib
from .params import (
    PIVParameters,
    _isize,
    _smarg,
    _step,
    get_vectorfield_size,
    mask_NSQECC,
    maxval,
    setPIVParameters,
)


def to_c_params(p: PIVParameters, corr_code: Optional[int] = None, mode: str = "auto",
                multipass_ok: bool = False) -> _lib.MqpivParams:
    """Resolve a PIVParameters into the C struct (step = interSize - overlap when
    ``step`` is ``nothing``, piv_parameters.jl:88).  ``mode`` picks the transform
    precision: "fp32", "fp64", "auto" (FP64 where FP32 misses the parity bar) or
    "mfma" (2-D DFTs as matrix-core products with hi + lo FP16 operands; single
    2-D calls, search sizes up to 96, no masked measure, no multipass).
    ``multipass`` travels separately (mqpiv_piv_multipass); callers that run a
    single pass leave ``multipass_ok`` False and reject multipass > 1."""
    if not (p.unpadded and p.good_pad):
        # piv_ops.jl:261-268: the padded branch prints an error and mis-indexes.
        raise NotImplementedError("only unpadded = good_pad = true is supported (as in the reference)")
    if p.multipass < 1:
        raise ValueError("multipass must be at least 1")
    if p.multipass != 1 and not multipass_ok:
        raise NotImplementedError("multipass > 1 runs through PIV / piv_device only")
    cp = _lib.MqpivParams()
    cp.corr_alg = p.corr_alg.code if corr_code is None else corr_code
    cp.ndims = p.ndims
    isz, smg, stp = _isize(p), _smarg(p), _step(p)
    for d in range(3):
        cp.inter_size[d] = isz[d] if d < p.ndims else 1
        cp.search_margin[d] = smg[d] if d < p.ndims else 0
        cp.step[d] = stp[d] if d < p.ndims else 1
    cp.compute_sn = 1 if p.computeSN else 0
    cp.precision_mode = _lib.PRECISION_MODES[mode]
    cp.threshold = p.threshold
    cp.mask_threshold = p.mask_threshold
    return cp


TRANSFORM_FORMS = {0: "f32", 1: "f64", 2: "split-f16", 3: "split-f16+f64", 4: "f32+f64"}


def transform_form(pivparams: PIVParameters, shape: Tuple[int, ...], mode: str = "auto",
                   masked: bool = False) -> Tuple[str, int]:
    """The transform form and length a call with these parameters (and a mask,
    if ``masked``) runs on an image of ``shape``: ("f32" | "f64" | "split-f16" |
    "split-f16+f64": the split form with its flagged IAs recomputed in FP64 |
    "f32+f64": FP32 transforms with their flagged IAs recomputed in FP64,
    length).  Host-only (mqpiv_transform_form); raises as the call would for an
    unsupported job."""
    p = copy.copy(pivparams)
    p.ndims = len(shape)
    cp = to_c_params(p, mask_NSQECC.code if masked else None, mode, multipass_ok=True)
    dims = (C.c_int64 * 3)(*(list(shape) + [1] * (3 - len(shape))))
    form, length = C.c_int32(0), C.c_int32(0)
    lib = _lib.load()
    rc = lib.mqpiv_transform_form(C.byref(cp), dims, C.byref(form), C.byref(length))
    if rc:
        raise _lib.MqpivError(rc, "no transform form for these parameters")
    return TRANSFORM_FORMS[form.value], length.value


def _as_input(a: Any) -> np.ndarray:
    a = np.asarray(a)
    if not np.issubdtype(a.dtype, np.number) or np.issubdtype(a.dtype, np.complexfloating):
        raise TypeError("PIV inputs must be real arrays")
    return a


def _common_dtype(*arrs: np.ndarray) -> np.dtype:
    if all(a.dtype == np.float32 for a in arrs):
        return np.dtype(np.float32)
    return np.dtype(np.float64)


def _ia_slices(vf_idx: Tuple[int, ...], p: PIVParameters) -> Tuple[slice, ...]:
    isz, stp = _isize(p), _step(p)
    return tuple(slice(v * s, v * s + i) for v, s, i in zip(vf_idx, stp, isz))


def host_skip_mask(img1: np.ndarray, mask: Optional[np.ndarray], p: PIVParameters,
                   masked: bool) -> Optional[np.ndarray]:
    """Per-IA skip flags for filters the device does not evaluate itself: a user
    ``filtFun``/``mask_filtFun`` other than ``maxval`` (piv_ops.jl:134-157).  In
    the double-mask form ``mask`` is mask_VF."""
    vf = get_vectorfield_size(img1.shape, p)
    n = int(np.prod(vf))
    out = np.zeros(n, dtype=np.uint8)
    for lin in range(n):
        idx = np.unravel_index(lin, vf, order="F")
        sl = _ia_slices(idx, p)
        if masked:
            skip = False
            if p.threshold > 0:
                skip = p.filtFun(img1[sl]) < p.threshold
            if p.mask_threshold > 0:
                skip = skip or (p.mask_filtFun(mask[sl]) < p.mask_threshold)
        else:
            skip = p.threshold >= 0 and p.filtFun(img1[sl]) < p.threshold
        out[lin] = 1 if skip else 0
    return out


def PIV(input1: Any, input2: Any, *args: Any, precision: Any = np.float64,  # noqa: N802
        device: int = 0, return_peaks: bool = False, mode: str = "auto"):
    """PIV(img1, img2[, pivparams]) / PIV(img1, img2, mask[, pivparams]) /
    PIV(img1, img2, mask_VF, mask_IA[, pivparams]) -> (VF, SN).

    With return_peaks=True also returns the 1-based integer correlation peaks,
    shape (N, nVF...) int32 (0 where no value beats 0 or the IA is skipped).
    ``mode``: transform precision, "auto" (default), "fp32", "fp64" or "mfma"."""
    args = list(args)
    pivparams = setPIVParameters()
    if args and isinstance(args[-1], PIVParameters):
        pivparams = args.pop()
    if len(args) > 2:
        raise TypeError("PIV takes at most two masks")
    img1 = _as_input(input1)
    img2 = _as_input(input2)
    masks = [_as_input(m) for m in args]
    # multi_quickPIV.jl:48/108/167
    if img1.shape != img2.shape or any(m.shape != img1.shape for m in masks):
        raise AssertionError("PIV inputs need to have the same size.")
    N = img1.ndim
    pivparams.ndims = N  # multi_quickPIV.jl:50 mutates the caller's parameters too
    masked = len(masks) > 0
    corr_code = mask_NSQECC.code if masked else None

    dt = _common_dtype(img1, img2, *masks)
    in_code = _lib.MQPIV_F32 if dt == np.float32 else _lib.MQPIV_F64
    a1 = np.asfortranarray(img1, dtype=dt)
    a2 = np.asfortranarray(img2, dtype=dt)
    m_ia = np.asfortranarray(masks[-1], dtype=dt) if masked else None
    m_vf = masks[0] if masked else None

    out_dt = np.dtype(precision)
    if out_dt not in (np.float32, np.float64):
        raise TypeError("precision must be Float32 or Float64")
    out_code = _lib.MQPIV_F32 if out_dt == np.float32 else _lib.MQPIV_F64

    vf = get_vectorfield_size(img1.shape, pivparams)
    VF = np.zeros((N, *vf), dtype=out_dt, order="F")
    SN = np.zeros(vf, dtype=out_dt, order="F") if pivparams.computeSN else None

    PK = np.zeros((N, *vf), dtype=np.int32, order="F") if return_peaks else None
    skip = None
    custom = (pivparams.filtFun is not maxval) or (masked and pivparams.mask_filtFun is not maxval)
    mp = int(pivparams.multipass)
    if mp > 1 and (custom or len(masks) == 2 or return_peaks):
        raise NotImplementedError("multipass > 1 takes the default filtFun, one mask and no peaks output")
    if custom:
        skip = host_skip_mask(img1, m_vf, pivparams, masked)
    if int(np.prod(vf)) == 0:
        return (VF, SN, PK) if return_peaks else (VF, SN)

    cp = to_c_params(pivparams, corr_code, mode, multipass_ok=True)
    dims = (C.c_int64 * 3)(*(list(img1.shape) + [1] * (3 - N)))
    ctx = _lib.context(device)
    if mp > 1:
        # multi_quickPIV.jl:57-88, coarse to fine (mqpiv_piv_multipass)
        rc = ctx.lib.mqpiv_piv_multipass(
            ctx.handle, C.byref(cp), mp, dims, in_code, a1.ctypes.data, a2.ctypes.data,
            m_ia.ctypes.data if m_ia is not None else None,
            out_code, VF.ctypes.data, SN.ctypes.data if SN is not None else None)
        ctx.check(rc)
        return VF, SN
    if len(masks) == 2 and skip is None:
        # multi_quickPIV.jl:155-213: skip filter on mask_VF, correlation on mask_IA;
        # the flags come from the device
        mv = np.asfortranarray(m_vf, dtype=dt)
        rc = ctx.lib.mqpiv_piv_masks(
            ctx.handle, C.byref(cp), dims, in_code, a1.ctypes.data, a2.ctypes.data, mv.ctypes.data,
            m_ia.ctypes.data, out_code, VF.ctypes.data, SN.ctypes.data if SN is not None else None,
            PK.ctypes.data if PK is not None else None)
        ctx.check(rc)
        return (VF, SN, PK) if return_peaks else (VF, SN)
    rc = ctx.lib.mqpiv_piv_ex(
        ctx.handle, C.byref(cp), dims, in_code,
        a1.ctypes.data, a2.ctypes.data,
        m_ia.ctypes.data if m_ia is not None else None,
        skip.ctypes.data if skip is not None else None,
        out_code, VF.ctypes.data, SN.ctypes.data if SN is not None else None,
        PK.ctypes.data if PK is not None else None,
    )
    ctx.check(rc)
    return (VF, SN, PK) if return_peaks else (VF, SN)


def piv_device(ctx: "_lib.Context", pivparams: PIVParameters, shape: Tuple[int, ...], in_dtype: int,
               d_img1: int, d_img2: int, d_mask: Optional[int], out_dtype: int, d_vf: int,
               d_sn: Optional[int], stream: Optional[int] = None, masked: bool = False,
               mode: str = "auto") -> None:
    """Asynchronous PIV over device-resident buffers (raw device pointers, e.g.
    ``tensor.data_ptr()``; tensors laid out column-major as in ``PIV``).
    ``pivparams.multipass`` > 1 runs the coarse-to-fine passes.  With
    ``stream`` None the work runs on the context's own stream, which is not
    ordered with torch's streams, so the device is synchronised first (the
    inputs and outputs may still be being written by torch)."""
    _order_before_library_stream(stream)
    cp = to_c_params(pivparams, mask_NSQECC.code if masked else None, mode, multipass_ok=True)
    dims = (C.c_int64 * 3)(*(list(shape) + [1] * (3 - len(shape))))
    if pivparams.multipass > 1:
        rc = ctx.lib.mqpiv_piv_multipass_device(ctx.handle, C.byref(cp), int(pivparams.multipass), dims, in_dtype,
                                                d_img1, d_img2, d_mask, out_dtype, d_vf, d_sn, stream)
    else:
        rc = ctx.lib.mqpiv_piv_device(ctx.handle, C.byref(cp), dims, in_dtype, d_img1, d_img2, d_mask,
                                      None, out_dtype, d_vf, d_sn, stream)
    ctx.check(rc)


class DeviceFrame:
    """A frame pair resident on the device, stored column-major (dim 1
    fastest) once.  Requests made from the same frame with ``DeviceRequest.
    from_frame`` share its buffers, so a batch builds the second image's
    integral arrays once for all of them."""

    def __init__(self, img1, img2):
        import torch

        if img1.shape != img2.shape:
            raise AssertionError("PIV inputs need to have the same size.")
        self.shape = tuple(img1.shape)
        dt = torch.float32 if (img1.dtype == torch.float32 and img2.dtype == torch.float32) else torch.float64
        perm = tuple(reversed(range(img1.dim())))
        self.dtype = dt
        self.a = img1.to(dt).permute(*perm).contiguous()
        self.b = img2.to(dt).permute(*perm).contiguous()
        self.in_code = _lib.MQPIV_F32 if dt == torch.float32 else _lib.MQPIV_F64


class DeviceRequest:
    """One frame pair resident on the device for batched / graph execution.

    ``img1``, ``img2`` (and ``mask``) are torch tensors of the frame's shape in
    the reference's axis order; they are stored column-major (dim 1 fastest)
    once, here (or taken from a ``DeviceFrame``, see ``from_frame``).  Outputs
    ``VF`` (N x nVF..., FP64) and ``SN`` are device tensors in the reference
    layout, filled by ``piv_batch`` / graph launches.
    """

    def __init__(self, img1, img2, pivparams: PIVParameters, mask=None, mode: str = "auto",
                 want_peaks: bool = False, frame: Optional[DeviceFrame] = None):
        import torch

        if frame is None:
            frame = DeviceFrame(img1, img2)
        if mask is not None and tuple(mask.shape) != frame.shape:
            raise AssertionError("PIV inputs need to have the same size.")
        N = len(frame.shape)
        pivparams.ndims = N
        self.shape = frame.shape
        perm = tuple(reversed(range(N)))
        self.a, self.b, self.in_code = frame.a, frame.b, frame.in_code
        self.m = mask.to(frame.dtype).permute(*perm).contiguous() if mask is not None else None
        self.vf_size = get_vectorfield_size(self.shape, pivparams)
        n = int(np.prod(self.vf_size))
        dev = frame.a.device
        self.VF = torch.zeros(n * N, dtype=torch.float64, device=dev)
        self.SN = torch.zeros(n, dtype=torch.float64, device=dev) if pivparams.computeSN else None
        self.PK = torch.zeros(n * N, dtype=torch.int32, device=dev) if want_peaks else None
        self.params = to_c_params(pivparams, mask_NSQECC.code if mask is not None else None, mode)

    @classmethod
    def from_frame(cls, frame: DeviceFrame, pivparams: PIVParameters, mask=None, mode: str = "auto",
                   want_peaks: bool = False) -> "DeviceRequest":
        return cls(None, None, pivparams, mask=mask, mode=mode, want_peaks=want_peaks, frame=frame)

    def c_request(self) -> "_lib.MqpivRequest":
        r = _lib.MqpivRequest()
        r.params = self.params
        for d in range(3):
            r.dims[d] = self.shape[d] if d < len(self.shape) else 1
        r.in_dtype = self.in_code
        r.out_dtype = _lib.MQPIV_F64
        r.img1, r.img2 = self.a.data_ptr(), self.b.data_ptr()
        r.mask = self.m.data_ptr() if self.m is not None else None
        r.skip_mask = None
        r.VF = self.VF.data_ptr()
        r.SN = self.SN.data_ptr() if self.SN is not None else None
        r.peaks = self.PK.data_ptr() if self.PK is not None else None
        return r

    def results(self):
        """(VF, SN[, peaks]) as numpy arrays in the reference layout (waits for
        all work on the device: the batch may run on a library-owned stream)."""
        import torch

        torch.cuda.synchronize(self.VF.device)
        N = len(self.shape)
        vf = tuple(self.vf_size)
        VF = self.VF.cpu().numpy().reshape((N, *vf), order="F")
        SN = self.SN.cpu().numpy().reshape(vf, order="F") if self.SN is not None else None
        if self.PK is not None:
            return VF, SN, self.PK.cpu().numpy().reshape((N, *vf), order="F")
        return VF, SN


def _order_before_library_stream(stream: Optional[int]) -> None:
    """The library's own stream (stream None) is a non-blocking stream: work
    queued by torch on the device (rendering inputs, zero-filling outputs) is
    not ordered before it, so wait for it."""
    if stream is None:
        import torch

        torch.cuda.synchronize()


def piv_batch(reqs, ctx=None, stream: Optional[int] = None) -> None:
    """All requests in one batched call (one kernel launch per algorithm /
    precision / transform-length group); asynchronous on ``stream`` (None: the
    context's stream, after a device synchronisation)."""
    _order_before_library_stream(stream)
    ctx = ctx or _lib.context(reqs[0].a.device.index or 0)
    arr = (_lib.MqpivRequest * len(reqs))(*[r.c_request() for r in reqs])
    ctx.check(ctx.lib.mqpiv_piv_batch_device(ctx.handle, arr, len(reqs), stream))


class BatchGraph:
    """A batch captured once into a HIP graph; ``launch`` replays it on the
    requests' current buffer contents."""

    def __init__(self, reqs, ctx=None):
        self.ctx = ctx or _lib.context(reqs[0].a.device.index or 0)
        self.reqs = list(reqs)  # keep the buffers alive
        arr = (_lib.MqpivRequest * len(reqs))(*[r.c_request() for r in reqs])
        h = C.c_void_p()
        self.ctx.check(self.ctx.lib.mqpiv_graph_create(self.ctx.handle, arr, len(reqs), C.byref(h)))
        self.handle = h

    def launch(self, stream: Optional[int] = None) -> None:
        _order_before_library_stream(stream)
        self.ctx.check(self.ctx.lib.mqpiv_graph_launch(self.handle, stream))

    def close(self) -> None:
        if self.handle:
            self.ctx.lib.mqpiv_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching request server (``mqpiv_server_*``): a native worker
    thread admits submitted pairs into execution slots by device footprint,
    runs each admitted set as one batched call and copies results back.

    ``submit`` takes host arrays in the reference's axis order; arrays that are
    already Fortran-ordered (for example views of pinned memory) are used in
    place and must not be modified until ``wait`` returns.  Outputs are
    allocated per request unless ``out=(VF, SN)`` supplies them."""

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 8 << 30, n_slots: int = 4,
                 max_batch: int = 64):
        self.lib = _lib.load()
        h = C.c_void_p()
        rc = self.lib.mqpiv_server_create(device, hbm_budget_bytes, n_slots, max_batch, C.byref(h))
        if rc != _lib.MQPIV_OK:
            raise _lib.MqpivError(rc, f"mqpiv_server_create(device={device}) failed")
        self.handle = h
        self._live: dict = {}

    def submit(self, img1: Any, img2: Any, pivparams: PIVParameters, mask: Any = None, mode: str = "auto",
               precision: Any = np.float64, out: Optional[Tuple[np.ndarray, Optional[np.ndarray]]] = None,
               return_peaks: bool = False) -> int:
        a1, a2 = _as_input(img1), _as_input(img2)
        if a1.shape != a2.shape or (mask is not None and np.shape(mask) != a1.shape):
            raise AssertionError("PIV inputs need to have the same size.")
        N = a1.ndim
        pivparams.ndims = N
        dt = _common_dtype(a1, a2, *([] if mask is None else [_as_input(mask)]))
        a1 = np.asfortranarray(a1, dtype=dt)
        a2 = np.asfortranarray(a2, dtype=dt)
        m = np.asfortranarray(_as_input(mask), dtype=dt) if mask is not None else None
        out_dt = np.dtype(precision)
        vf = get_vectorfield_size(a1.shape, pivparams)
        if out is None:
            VF = np.zeros((N, *vf), dtype=out_dt, order="F")
            SN = np.zeros(vf, dtype=out_dt, order="F") if pivparams.computeSN else None
        else:
            VF, SN = out
            out_dt = VF.dtype
        PK = np.zeros((N, *vf), dtype=np.int32, order="F") if return_peaks else None
        r = _lib.MqpivRequest()
        r.params = to_c_params(pivparams, mask_NSQECC.code if mask is not None else None, mode)
        for d in range(3):
            r.dims[d] = a1.shape[d] if d < N else 1
        r.in_dtype = _lib.MQPIV_F32 if dt == np.float32 else _lib.MQPIV_F64
        r.out_dtype = _lib.MQPIV_F32 if out_dt == np.float32 else _lib.MQPIV_F64
        r.img1, r.img2 = a1.ctypes.data, a2.ctypes.data
        r.mask = m.ctypes.data if m is not None else None
        r.skip_mask = None
        r.VF = VF.ctypes.data
        r.SN = SN.ctypes.data if SN is not None else None
        r.peaks = PK.ctypes.data if PK is not None else None
        t = C.c_uint64()
        rc = self.lib.mqpiv_server_submit(self.handle, C.byref(r), C.byref(t))
        if rc != _lib.MQPIV_OK:
            raise _lib.MqpivError(rc, "mqpiv_server_submit rejected the request")
        self._live[t.value] = (a1, a2, m, VF, SN, PK)
        return t.value

    def submit_device(self, dreq: "DeviceRequest") -> int:
        """Submit a pair resident in device memory (a DeviceRequest, or any
        object whose ``c_request()`` names device buffers); its outputs stay on
        the device."""
        r = dreq.c_request()
        t = C.c_uint64()
        rc = self.lib.mqpiv_server_submit_device(self.handle, C.byref(r), C.byref(t))
        if rc != _lib.MQPIV_OK:
            raise _lib.MqpivError(rc, "mqpiv_server_submit_device rejected the request")
        self._live[t.value] = dreq
        return t.value

    def wait(self, ticket: int, timeout: float = -1.0):
        """(VF, SN[, peaks]) of a finished host request, or the DeviceRequest of
        a device one; None if ``timeout`` expired."""
        st = C.c_int32()
        rc = self.lib.mqpiv_server_wait(self.handle, ticket, timeout, C.byref(st))
        if rc != _lib.MQPIV_OK:
            raise _lib.MqpivError(rc, f"unknown ticket {ticket}")
        if st.value == -1:
            return None
        live = self._live.pop(ticket)
        if st.value != _lib.MQPIV_OK:
            msg = self.lib.mqpiv_server_last_error(self.handle).decode(errors="replace")
            raise _lib.MqpivError(st.value, f"request {ticket} failed: {msg}")
        if not isinstance(live, tuple):  # a device request: outputs stay on the device
            return live
        _, _, _, VF, SN, PK = live
        return (VF, SN, PK) if PK is not None else (VF, SN)

    def stats(self) -> dict:
        s = _lib.MqpivServerStats()
        rc = self.lib.mqpiv_server_stats_get(self.handle, C.byref(s))
        if rc != _lib.MQPIV_OK:
            raise _lib.MqpivError(rc, "mqpiv_server_stats_get failed")
        return {k: getattr(s, k) for k, _ in s._fields_}

    def reset_stats(self) -> None:
        self.lib.mqpiv_server_reset_stats(self.handle)

    def close(self) -> None:
        if self.handle:
            self.lib.mqpiv_server_destroy(self.handle)
            self.handle = None
            self._live.clear()

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass
