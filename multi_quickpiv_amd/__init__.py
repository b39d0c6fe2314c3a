"""multi_quickpiv_amd -- MI355X-native cross-correlation PIV.

The drop-in surface of multi_quickPIV's hot path:

    from multi_quickpiv_amd import PIV, setPIVParameters
    VF, SN = PIV(img1, img2, setPIVParameters(corr_alg="nsqecc", interSize=32,
                                              searchMargin=16, step=16, computeSN=True))

Computation runs in hand-written gfx950 HIP kernels behind the C ABI of
include/mqpiv.h (libmqpiv.so, built in-tree).
"""
from .params import (  # noqa: F401
    CORRTYPES,
    FFTCC,
    NSQECC,
    ZNCC,
    PIVParameters,
    _isize,
    _ovp,
    _smarg,
    _ssize,
    _step,
    get_vectorfield_size,
    mask_NSQECC,
    maxval,
    parseCorr,
    setPIVParameters,
    toDims3,
)
from .piv import PIV, BatchGraph, DeviceFrame, DeviceRequest, Server, piv_batch, piv_device  # noqa: F401

__all__ = [
    "PIV",
    "piv_device",
    "piv_batch",
    "DeviceFrame",
    "DeviceRequest",
    "BatchGraph",
    "Server",
    "setPIVParameters",
    "PIVParameters",
    "FFTCC",
    "ZNCC",
    "NSQECC",
    "mask_NSQECC",
    "parseCorr",
]
