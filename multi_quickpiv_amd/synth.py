"""Synthetic particle image pairs for tests and benchmarks.

Particle model: I(x) = bg + sum_p A_p exp(-|x - c_p|^2 / (2 sigma^2)), rendered
analytically and truncated at 4 sigma, A_p ~ U[0.5, 1.0], centres c_p uniform over
the image.  Frame 2 renders the same particles at c_p + d(c_p) (exact sub-pixel
shifts, no image warping).  Each frame gets independent N(0, noise) noise.
Axis order is the reference's: axis 0 = dim 1 (rows, U), axis 1 = dim 2 (V),
axis 2 = dim 3 (W); pixel i has its centre at coordinate i.  Seeds drive a NumPy
PCG64 generator.
"""
from __future__ import annotations

import math
from typing import Callable, Optional, Sequence, Tuple

import numpy as np
import torch

Displacement = Callable[[np.ndarray], np.ndarray]  # (n, N) centres -> (n, N) shifts


def uniform(d: Sequence[float]) -> Displacement:
    dv = np.asarray(d, dtype=np.float64)
    return lambda c: np.broadcast_to(dv, c.shape).copy()


def lamb_oseen(shape: Sequence[int], core_radius: float = 300.0, peak: float = 8.0,
               drift: Sequence[float] = (1.5, 0.5)) -> Displacement:
    """Lamb-Oseen vortex centred on the image, u_theta(r) = G/(2 pi r)(1 - exp(-r^2/rc^2)),
    G scaled so that max u_theta = peak, plus a uniform drift."""
    ctr = np.array([(s - 1) / 2.0 for s in shape[:2]])
    r = np.linspace(1e-3, 5 * core_radius, 200001)
    fmax = np.max((1 - np.exp(-(r ** 2) / core_radius ** 2)) / r)
    k = peak / fmax  # = G / (2 pi)
    dr = np.asarray(drift, dtype=np.float64)

    def f(c: np.ndarray) -> np.ndarray:
        x = c[:, 0] - ctr[0]
        y = c[:, 1] - ctr[1]
        rr = np.sqrt(x * x + y * y)
        rr = np.maximum(rr, 1e-9)
        ut = k * (1 - np.exp(-(rr ** 2) / core_radius ** 2)) / rr
        d = np.empty_like(c)
        d[:, 0] = -ut * y / rr + dr[0]
        d[:, 1] = ut * x / rr + dr[1]
        return d

    return f


def shear3d(rate: float = 0.01, drift: Sequence[float] = (2.3, -1.1, 0.7)) -> Displacement:
    """u = rate * z + drift (z = 0-based dim-3 coordinate), v, w uniform."""
    dr = np.asarray(drift, dtype=np.float64)

    def f(c: np.ndarray) -> np.ndarray:
        d = np.broadcast_to(dr, c.shape).copy()
        d[:, 0] += rate * c[:, 2]
        return d

    return f


def render(shape: Sequence[int], centres: np.ndarray, amps: np.ndarray, sigma: float, bg: float,
           device: str = "cpu", dtype=torch.float32) -> torch.Tensor:
    """Sum of truncated Gaussians (|x - c| <= 4 sigma) on a zero-initialised grid plus bg."""
    nd = len(shape)
    img = torch.zeros(int(np.prod(shape)), dtype=torch.float64, device=device)
    c = torch.as_tensor(centres, dtype=torch.float64, device=device)
    a = torch.as_tensor(amps, dtype=torch.float64, device=device)
    r = int(math.ceil(4 * sigma))
    base = torch.round(c).to(torch.int64)
    inv2s2 = 1.0 / (2 * sigma * sigma)
    lim2 = (4 * sigma) ** 2
    offs = range(-r, r + 1)
    # strides for Fortran (dim 1 fastest) linear indices
    strides = [1] * nd
    for d in range(1, nd):
        strides[d] = strides[d - 1] * shape[d - 1]
    outer = [(o,) for o in offs] if nd == 2 else [(o1, o2) for o1 in offs for o2 in offs]
    inner = torch.arange(-r, r + 1, device=device, dtype=torch.int64)
    for oo in outer:
        # axis 0 varies in `inner`; the remaining axes are fixed by `oo`
        pos = [base[:, 0:1] + inner[None, :]]
        for d, o in enumerate(oo, start=1):
            pos.append((base[:, d] + o)[:, None].expand(-1, inner.numel()))
        d2 = torch.zeros_like(pos[0], dtype=torch.float64)
        ok = torch.ones_like(pos[0], dtype=torch.bool)
        lin = torch.zeros_like(pos[0])
        for d in range(nd):
            diff = pos[d].to(torch.float64) - c[:, d:d + 1]
            d2 = d2 + diff * diff
            ok &= (pos[d] >= 0) & (pos[d] < shape[d])
            lin = lin + pos[d].clamp(0, shape[d] - 1) * strides[d]
        ok &= d2 <= lim2
        w = a[:, None] * torch.exp(-d2 * inv2s2)
        img.index_add_(0, lin[ok], w[ok])
    img += bg
    # back to array order: Fortran linear index -> tensor of `shape` (axis 0 fastest)
    return img.reshape(tuple(reversed(shape))).permute(*reversed(range(nd))).to(dtype)


def particle_pair(shape: Sequence[int], n_particles: int, sigma: float, disp: Displacement,
                  seed: int, bg: float = 0.1, noise: float = 0.02, device: str = "cpu",
                  dtype=torch.float32) -> Tuple[torch.Tensor, torch.Tensor]:
    """Two frames of the particle model; tensors of `shape` (axis 0 = dim 1)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    nd = len(shape)
    c = np.stack([rng.uniform(-0.5, s - 0.5, n_particles) for s in shape], axis=1)
    amp = rng.uniform(0.5, 1.0, n_particles)
    c2 = c + disp(c)
    im1 = render(shape, c, amp, sigma, bg, device, torch.float64)
    im2 = render(shape, c2, amp, sigma, bg, device, torch.float64)
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    im1 += noise * torch.randn(im1.shape, generator=g, device=device, dtype=torch.float64)
    im2 += noise * torch.randn(im2.shape, generator=g, device=device, dtype=torch.float64)
    return im1.to(dtype), im2.to(dtype)


def ellipse_mask(shape: Sequence[int], n: int = 20, cover: float = 0.3, seed: int = 0,
                 device: str = "cpu", dtype=torch.float32) -> torch.Tensor:
    """Binary mask, 0 inside n seeded random ellipses (about `cover` of the area)
    and 1 elsewhere."""
    rng = np.random.Generator(np.random.PCG64(seed))
    H, W = shape
    area = cover * H * W / n
    y = torch.arange(H, device=device, dtype=torch.float64)[:, None]
    x = torch.arange(W, device=device, dtype=torch.float64)[None, :]
    m = torch.ones((H, W), dtype=torch.bool, device=device)
    for _ in range(n):
        cy, cx = rng.uniform(0, H), rng.uniform(0, W)
        ratio = rng.uniform(0.4, 1.0)
        a = math.sqrt(area / (math.pi * ratio))
        b = a * ratio
        th = rng.uniform(0, math.pi)
        ct, st = math.cos(th), math.sin(th)
        u = (y - cy) * ct + (x - cx) * st
        v = -(y - cy) * st + (x - cx) * ct
        m &= ~((u / a) ** 2 + (v / b) ** 2 <= 1.0)
    return m.to(dtype)


def n_particles(shape: Sequence[int], density: float) -> int:
    return int(round(density * float(np.prod(shape))))


def config_c1(seed: int = 1000) -> Tuple[np.ndarray, np.ndarray]:
    """256x256 Float64 pair: 2000 particles, sigma 1.2, shift (3.4, -2.7)."""
    a, b = particle_pair((256, 256), 2000, 1.2, uniform((3.4, -2.7)), seed, dtype=torch.float64)
    return a.numpy(), b.numpy()


def config_c2(size: int = 2048, seed: int = 2000, device: str = "cpu"):
    """FP32 2048^2 pair at 0.05 particles/px, Lamb-Oseen vortex + drift."""
    shape = (size, size)
    return particle_pair(shape, n_particles(shape, 0.05), 1.2, lamb_oseen(shape), seed,
                         device=device, dtype=torch.float32)


def config_c3(shape=(512, 512, 256), seed: int = 3000, device: str = "cpu"):
    """FP32 volume pair at 0.01 particles/voxel, sigma 1.5, shear + drift."""
    return particle_pair(shape, n_particles(shape, 0.01), 1.5, shear3d(), seed, device=device,
                         dtype=torch.float32)
