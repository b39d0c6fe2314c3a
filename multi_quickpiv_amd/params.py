"""PIV parameters: the configuration surface of multi_quickPIV.

Mirrors src/piv_parameters.jl:1-88 of the reference: the algorithm tag types,
``parseCorr``, the ``PIVParameters`` record, ``setPIVParameters`` with the same
keyword defaults, and the ``_isize/_smarg/_ssize/_step/_ovp`` accessors.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Any, Callable, Optional, Tuple

import numpy as np


# --- algorithm tags (piv_parameters.jl:2-7) ---------------------------------
class _CorrTag:
    code: int = -1
    name: str = ""

    def __eq__(self, other: object) -> bool:
        return type(self) is type(other)

    def __hash__(self) -> int:
        return hash(type(self))

    def __repr__(self) -> str:
        return f"{self.name}()"


class FFTCC(_CorrTag):
    code, name = 0, "FFTCC"


class ZNCC(_CorrTag):
    code, name = 1, "ZNCC"


class NSQECC(_CorrTag):
    code, name = 2, "NSQECC"


class mask_NSQECC(_CorrTag):  # noqa: N801 (reference name)
    code, name = 3, "mask_NSQECC"


CORRTYPES = (FFTCC, ZNCC, NSQECC, mask_NSQECC)


def maxval(a: np.ndarray) -> float:
    """Default background filter: the maximum of the window (common_ops.jl:29-47)."""
    return float(np.max(a))


def toDims3(x: Any) -> Optional[Tuple[int, int, int]]:  # noqa: N802 (reference name)
    """piv_parameters.jl:9-12."""
    if x is None:
        return None
    if isinstance(x, (int, np.integer)) and not isinstance(x, bool):
        return (int(x), int(x), int(x))
    if isinstance(x, tuple) and len(x) == 2:
        return (int(x[0]), int(x[1]), 0)
    if isinstance(x, tuple) and len(x) == 3:
        return (int(x[0]), int(x[1]), int(x[2]))
    return None


def parseCorr(corr: str) -> _CorrTag:  # noqa: N802 (reference name)
    """piv_parameters.jl:14-23: unknown names fall back to NSQECC; the three
    truncated names name types the reference never defines (UndefVarError)."""
    c = corr.lower()
    if c == "zncc":
        return ZNCC()
    if c == "fft":
        return FFTCC()
    if c == "nsqecc":
        return NSQECC()
    if c in ("znc", "fftc", "nsqec"):
        raise NameError(f"UndefVarError: {c.upper()} not defined")
    return NSQECC()


@dataclass
class PIVParameters:
    """piv_parameters.jl:29-46 (field for field)."""

    corr_alg: _CorrTag
    unpadded: bool
    good_pad: bool
    odd_pad: bool
    interSize: Tuple[int, int, int]
    searchMargin: Tuple[int, int, int]
    overlap: Tuple[int, int, int]
    step: Optional[Tuple[int, int, int]]
    multipass: int
    computeSN: bool
    filtFun: Callable[[np.ndarray], float]
    threshold: float
    mask_filtFun: Callable[[np.ndarray], float]
    mask_threshold: float
    ndims: int
    ovp_th: float


def setPIVParameters(  # noqa: N802 (reference name)
    *,
    corr_alg: str = "nsqecc",
    unpadded: bool = True,
    good_pad: bool = True,
    odd_pad: bool = False,
    interSize: Any = 32,
    searchMargin: Any = 10,
    overlap: Any = 0,
    step: Any = None,
    multipass: int = 1,
    computeSN: bool = False,
    filtFun: Optional[Callable] = None,
    threshold: float = -1.0,
    mask_filtFun: Optional[Callable] = None,
    mask_threshold: float = -1.0,
    ndims: int = 3,
    ovp_th: float = 0.5,
) -> PIVParameters:
    """piv_parameters.jl:68-82 defaults.  The outer constructor of the reference
    (:49-65) passes ``good_pad`` into the ``unpadded`` slot and vice versa; that
    swap is kept (it only matters when the two differ)."""
    return PIVParameters(
        corr_alg=parseCorr(corr_alg),
        unpadded=bool(good_pad),
        good_pad=bool(unpadded),
        odd_pad=bool(odd_pad),
        interSize=toDims3(interSize),
        searchMargin=toDims3(searchMargin),
        overlap=toDims3(overlap),
        step=toDims3(step),
        multipass=int(multipass),
        computeSN=bool(computeSN),
        filtFun=filtFun if filtFun is not None else maxval,
        threshold=float(threshold),
        mask_filtFun=mask_filtFun if mask_filtFun is not None else maxval,
        mask_threshold=float(mask_threshold),
        ndims=int(ndims),
        ovp_th=float(ovp_th),
    )


# --- accessors (piv_parameters.jl:84-88) -------------------------------------
def _ovp(p: PIVParameters, scale: int = 1) -> Tuple[int, ...]:
    return tuple(v * scale for v in p.overlap[: p.ndims])


def _isize(p: PIVParameters, scale: int = 1) -> Tuple[int, ...]:
    return tuple(v * scale for v in p.interSize[: p.ndims])


def _smarg(p: PIVParameters, scale: int = 1) -> Tuple[int, ...]:
    return tuple(v * scale for v in p.searchMargin[: p.ndims])


def _ssize(p: PIVParameters, scale: int = 1) -> Tuple[int, ...]:
    return tuple(i + 2 * m for i, m in zip(_isize(p, scale), _smarg(p, scale)))


def _step(p: PIVParameters, scale: int = 1) -> Tuple[int, ...]:
    if p.step is None:
        return tuple(i - o for i, o in zip(_isize(p, scale), _ovp(p, scale)))
    return tuple(v * scale for v in p.step[: p.ndims])


def get_vectorfield_size(input_size: Tuple[int, ...], p: PIVParameters) -> Tuple[int, ...]:
    """piv_ops.jl:6-14: length(1:step:(dim - interSize + 1)) per dim."""
    out = []
    for dim, i, s in zip(input_size, _isize(p), _step(p)):
        stop = dim - i + 1
        out.append((stop - 1) // s + 1 if stop >= 1 else 0)
    return tuple(out)
