"""ctypes binding of libmqpiv.so (the C ABI of include/mqpiv.h).

The library is loaded from the package directory.  There is no fallback: if it
is missing, importing the PIV entry points raises, so a GPU run can never
silently take another path.
"""
from __future__ import annotations

import ctypes as C
import os
import threading
from pathlib import Path

# The in-tree build; MQPIV_LIB names another build of the same library (for
# example one compiled with different tuning flags).
LIB_PATH = Path(os.environ.get("MQPIV_LIB") or (Path(__file__).resolve().parent / "libmqpiv.so"))

MQPIV_OK = 0
MQPIV_ERR_SIZE_MISMATCH = 1
MQPIV_ERR_BAD_PARAMS = 2
MQPIV_ERR_UNSUPPORTED = 3
MQPIV_ERR_OOM = 4
MQPIV_ERR_DEVICE = 5
MQPIV_ERR_NULL = 6

MQPIV_F32 = 0
MQPIV_F64 = 1

MQPIV_CMP_GT = 0
MQPIV_CMP_LT = 1
MQPIV_REPLACE_ZERO = 0
MQPIV_REPLACE_MEDIAN = 1

MQPIV_PLAIN_FP32 = 0
MQPIV_SPLIT_MFMA = 1
MQPIV_FP64 = 2
MQPIV_AUTO = 3
PRECISION_MODES = {"fp32": MQPIV_PLAIN_FP32, "mfma": MQPIV_SPLIT_MFMA, "fp64": MQPIV_FP64, "auto": MQPIV_AUTO}

# Every symbol include/mqpiv.h declares.
EXPORTS = (
    "mqpiv_ctx_create",
    "mqpiv_ctx_destroy",
    "mqpiv_last_error",
    "mqpiv_vf_size",
    "mqpiv_transform_form",
    "mqpiv_piv",
    "mqpiv_piv_ex",
    "mqpiv_piv_device",
    "mqpiv_piv_device_noalloc",
    "mqpiv_piv_masks",
    "mqpiv_piv_masks_device",
    "mqpiv_piv_multipass",
    "mqpiv_piv_multipass_device",
    "mqpiv_version",
    "mqpiv_profile_begin",
    "mqpiv_profile_end",
    "mqpiv_piv_batch_device",
    "mqpiv_graph_create",
    "mqpiv_graph_launch",
    "mqpiv_graph_destroy",
    "mqpiv_server_create",
    "mqpiv_server_submit",
    "mqpiv_server_submit_device",
    "mqpiv_server_wait",
    "mqpiv_server_stats_get",
    "mqpiv_server_reset_stats",
    "mqpiv_server_last_error",
    "mqpiv_server_destroy",
    "mqpiv_pp_magnitudes",
    "mqpiv_pp_std_mask",
    "mqpiv_pp_th_mask",
    "mqpiv_pp_replace",
    "mqpiv_pp_average",
    "mqpiv_pp_collectiveness",
    "mqpiv_pp_similarity_averaging",
    "mqpiv_pp_local_magnitude_mask",
)


class MqpivParams(C.Structure):
    _fields_ = [
        ("corr_alg", C.c_int32),
        ("ndims", C.c_int32),
        ("inter_size", C.c_int64 * 3),
        ("search_margin", C.c_int64 * 3),
        ("step", C.c_int64 * 3),
        ("compute_sn", C.c_int32),
        ("precision_mode", C.c_int32),
        ("threshold", C.c_double),
        ("mask_threshold", C.c_double),
    ]


class MqpivRequest(C.Structure):
    _fields_ = [
        ("params", MqpivParams),
        ("dims", C.c_int64 * 3),
        ("in_dtype", C.c_int32),
        ("out_dtype", C.c_int32),
        ("img1", C.c_void_p),
        ("img2", C.c_void_p),
        ("mask", C.c_void_p),
        ("skip_mask", C.c_void_p),
        ("VF", C.c_void_p),
        ("SN", C.c_void_p),
        ("peaks", C.c_void_p),
    ]


class MqpivVfDesc(C.Structure):
    _fields_ = [("ndims", C.c_int32), ("dtype", C.c_int32), ("size", C.c_int64 * 3)]


class MqpivServerStats(C.Structure):
    _fields_ = [
        ("completed", C.c_uint64),
        ("ia_done", C.c_int64),
        ("pending", C.c_uint64),
        ("p50_ms", C.c_double),
        ("p99_ms", C.c_double),
        ("max_ms", C.c_double),
        ("pairs_per_s", C.c_double),
        ("ia_per_s", C.c_double),
    ]


class MqpivError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"mqpiv error {code}: {msg}")
        self.code = code


_lib = None
_lock = threading.Lock()


def load() -> C.CDLL:
    """Load libmqpiv.so (raises OSError when it has not been built)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not LIB_PATH.exists():
            raise OSError(
                f"{LIB_PATH} is missing: build it with `python -m multi_quickpiv_amd.build`"
            )
        lib = C.CDLL(str(LIB_PATH))
        vp = C.c_void_p
        lib.mqpiv_version.restype = C.c_char_p
        lib.mqpiv_version.argtypes = []
        lib.mqpiv_ctx_create.restype = C.c_int
        lib.mqpiv_ctx_create.argtypes = [C.c_int, C.c_uint64, C.POINTER(vp)]
        lib.mqpiv_ctx_destroy.restype = C.c_int
        lib.mqpiv_ctx_destroy.argtypes = [vp]
        lib.mqpiv_last_error.restype = C.c_char_p
        lib.mqpiv_last_error.argtypes = [vp]
        lib.mqpiv_vf_size.restype = C.c_int
        lib.mqpiv_vf_size.argtypes = [C.POINTER(MqpivParams), C.POINTER(C.c_int64),
                                      C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        lib.mqpiv_transform_form.restype = C.c_int
        lib.mqpiv_transform_form.argtypes = [C.POINTER(MqpivParams), C.POINTER(C.c_int64),
                                             C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
        lib.mqpiv_piv.restype = C.c_int
        lib.mqpiv_piv.argtypes = [vp, C.POINTER(MqpivParams), C.POINTER(C.c_int64), C.c_int32,
                                  vp, vp, vp, vp, C.c_int32, vp, vp]
        lib.mqpiv_piv_ex.restype = C.c_int
        lib.mqpiv_piv_ex.argtypes = [vp, C.POINTER(MqpivParams), C.POINTER(C.c_int64), C.c_int32,
                                     vp, vp, vp, vp, C.c_int32, vp, vp, vp]
        lib.mqpiv_piv_device.restype = C.c_int
        lib.mqpiv_piv_device.argtypes = [vp, C.POINTER(MqpivParams), C.POINTER(C.c_int64), C.c_int32,
                                         vp, vp, vp, vp, C.c_int32, vp, vp, vp]
        lib.mqpiv_piv_device_noalloc.restype = C.c_int
        lib.mqpiv_piv_device_noalloc.argtypes = [vp, C.POINTER(MqpivParams), C.POINTER(C.c_int64), C.c_int32,
                                                 vp, vp, vp, vp, C.c_int32, vp, vp, vp]
        lib.mqpiv_piv_masks.restype = C.c_int
        lib.mqpiv_piv_masks.argtypes = [vp, C.POINTER(MqpivParams), C.POINTER(C.c_int64), C.c_int32,
                                        vp, vp, vp, vp, C.c_int32, vp, vp, vp]
        lib.mqpiv_piv_masks_device.restype = C.c_int
        lib.mqpiv_piv_masks_device.argtypes = [vp, C.POINTER(MqpivParams), C.POINTER(C.c_int64), C.c_int32,
                                               vp, vp, vp, vp, C.c_int32, vp, vp, vp]
        lib.mqpiv_piv_multipass.restype = C.c_int
        lib.mqpiv_piv_multipass.argtypes = [vp, C.POINTER(MqpivParams), C.c_int32, C.POINTER(C.c_int64),
                                            C.c_int32, vp, vp, vp, C.c_int32, vp, vp]
        lib.mqpiv_piv_multipass_device.restype = C.c_int
        lib.mqpiv_piv_multipass_device.argtypes = [vp, C.POINTER(MqpivParams), C.c_int32, C.POINTER(C.c_int64),
                                                   C.c_int32, vp, vp, vp, C.c_int32, vp, vp, vp]
        lib.mqpiv_piv_batch_device.restype = C.c_int
        lib.mqpiv_piv_batch_device.argtypes = [vp, C.POINTER(MqpivRequest), C.c_int32, vp]
        lib.mqpiv_graph_create.restype = C.c_int
        lib.mqpiv_graph_create.argtypes = [vp, C.POINTER(MqpivRequest), C.c_int32, C.POINTER(vp)]
        lib.mqpiv_graph_launch.restype = C.c_int
        lib.mqpiv_graph_launch.argtypes = [vp, vp]
        lib.mqpiv_graph_destroy.restype = C.c_int
        lib.mqpiv_graph_destroy.argtypes = [vp]
        lib.mqpiv_profile_begin.restype = C.c_int
        lib.mqpiv_profile_begin.argtypes = [vp]
        lib.mqpiv_profile_end.restype = C.c_int
        lib.mqpiv_profile_end.argtypes = [vp, C.POINTER(C.c_double), C.c_int32, C.POINTER(C.c_int32)]
        lib.mqpiv_server_create.restype = C.c_int
        lib.mqpiv_server_create.argtypes = [C.c_int, C.c_uint64, C.c_int32, C.c_int32, C.POINTER(vp)]
        lib.mqpiv_server_submit.restype = C.c_int
        lib.mqpiv_server_submit.argtypes = [vp, C.POINTER(MqpivRequest), C.POINTER(C.c_uint64)]
        lib.mqpiv_server_submit_device.restype = C.c_int
        lib.mqpiv_server_submit_device.argtypes = [vp, C.POINTER(MqpivRequest), C.POINTER(C.c_uint64)]
        lib.mqpiv_server_wait.restype = C.c_int
        lib.mqpiv_server_wait.argtypes = [vp, C.c_uint64, C.c_double, C.POINTER(C.c_int32)]
        lib.mqpiv_server_stats_get.restype = C.c_int
        lib.mqpiv_server_stats_get.argtypes = [vp, C.POINTER(MqpivServerStats)]
        lib.mqpiv_server_last_error.restype = C.c_char_p
        lib.mqpiv_server_last_error.argtypes = [vp]
        lib.mqpiv_server_reset_stats.restype = C.c_int
        lib.mqpiv_server_reset_stats.argtypes = [vp]
        lib.mqpiv_server_destroy.restype = C.c_int
        lib.mqpiv_server_destroy.argtypes = [vp]
        D = C.POINTER(MqpivVfDesc)
        R = C.POINTER(C.c_int64)
        for name, args in (
            ("mqpiv_pp_magnitudes", [vp, D, vp, vp, vp]),
            ("mqpiv_pp_std_mask", [vp, D, vp, C.c_double, vp, vp]),
            ("mqpiv_pp_th_mask", [vp, D, vp, C.c_double, C.c_int32, vp, vp]),
            ("mqpiv_pp_replace", [vp, D, vp, vp, C.c_int32, R, vp, vp]),
            ("mqpiv_pp_average", [vp, D, vp, R, vp, vp]),
            ("mqpiv_pp_collectiveness", [vp, D, vp, R, C.c_double, vp, vp]),
            ("mqpiv_pp_similarity_averaging", [vp, D, vp, R, C.c_double, C.c_int32, C.c_int32, C.c_int32, vp,
                                               vp]),
            ("mqpiv_pp_local_magnitude_mask", [vp, D, vp, C.c_double, R, vp, vp, vp]),
        ):
            f = getattr(lib, name)
            f.restype = C.c_int
            f.argtypes = args
        _lib = lib
        return lib


class Context:
    """One libmqpiv context (a HIP stream plus device staging) on one device."""

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 0):
        self.lib = load()
        h = C.c_void_p()
        rc = self.lib.mqpiv_ctx_create(device, hbm_budget_bytes, C.byref(h))
        if rc != MQPIV_OK:
            raise MqpivError(rc, f"mqpiv_ctx_create(device={device}) failed")
        self.handle = h
        self.device = device

    def check(self, rc: int) -> None:
        if rc != MQPIV_OK:
            msg = self.lib.mqpiv_last_error(self.handle).decode(errors="replace")
            raise MqpivError(rc, msg)

    def profile_begin(self) -> None:
        self.check(self.lib.mqpiv_profile_begin(self.handle))

    def profile_end(self, cap: int = 1 << 16) -> list:
        """Durations (ms) of the correlation kernels launched since profile_begin."""
        buf = (C.c_double * cap)()
        n = C.c_int32()
        self.check(self.lib.mqpiv_profile_end(self.handle, buf, cap, C.byref(n)))
        return [buf[i] for i in range(min(n.value, cap))]

    def close(self) -> None:
        if self.handle:
            self.lib.mqpiv_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self) -> None:  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass


_ctx: dict[int, Context] = {}


def context(device: int = 0) -> Context:
    with _lock:
        c = _ctx.get(device)
    if c is None:
        c = Context(device)
        with _lock:
            _ctx[device] = c
    return c
