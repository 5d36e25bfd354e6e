"""Synthetic porous-media geometries (uint8, shape (nz, ny, nx), 0 = fluid,
1 = solid, x fastest when written with ``tofile`` -- the reference's .raw
layout, README.md "Input data").

* ``tube``: straight cylindrical pore along z (Hagen-Poiseuille case).
* ``sphere_pack``: solid balls with radii uniform in [rmin, rmax] and centres
  uniform in the box from a seed, inserted until the porosity reaches the
  target (non-periodic clipping).
* ``gaussian_field``: thresholded Gaussian random field -- seeded white noise
  filtered in Fourier space to the covariance exp(-r^2/l^2) (amplitude filter
  exp(-k^2 l^2 / 8)) and thresholded at the porosity quantile.
* ``percolates``: a 6-connected fluid component touching both z faces (the
  connectivity check of the reference README, scipy.ndimage.label).
"""
from __future__ import annotations

import numpy as np


def tube(nx: int, ny: int, nz: int, cx: float, cy: float, r: float) -> np.ndarray:
    """Fluid iff (x-cx)^2 + (y-cy)^2 <= r^2 over integer voxel indices."""
    x = np.arange(nx)[None, :]
    y = np.arange(ny)[:, None]
    fl = (x - cx) ** 2 + (y - cy) ** 2 <= r * r
    g = np.ones((nz, ny, nx), np.uint8)
    g[:, fl] = 0
    return g


def channel(nx: int, ny: int, nz: int, wall: int = 2) -> np.ndarray:
    """Rectangular duct along z with solid walls of the given thickness."""
    g = np.ones((nz, ny, nx), np.uint8)
    g[:, wall:ny - wall, wall:nx - wall] = 0
    return g


def percolates(g: np.ndarray, axis: int = 0) -> bool:
    """True if a 6-connected fluid cluster touches both faces normal to
    ``axis`` (axis 0 = z for (nz, ny, nx) arrays)."""
    from scipy import ndimage

    lab, n = ndimage.label(g == 0)
    if n == 0:
        return False
    a = np.moveaxis(lab, axis, 0)
    lo = np.unique(a[0])
    hi = np.unique(a[-1])
    common = np.intersect1d(lo[lo > 0], hi[hi > 0])
    return common.size > 0


def sphere_pack(n, phi: float, rmin: float, rmax: float, seed: int,
                require_percolation: bool = True, max_tries: int = 50, axes=(0,)) -> np.ndarray:
    """Overlapping solid spheres until porosity <= phi; ``n`` is an int (cube)
    or an (nz, ny, nx) tuple.  Re-seeds (seed + 1000*try) until the fluid
    percolates along every array axis in ``axes`` (0 = z)."""
    shape = (n, n, n) if np.isscalar(n) else tuple(n)
    for attempt in range(max_tries):
        rng = np.random.default_rng(seed + 1000 * attempt)
        g = np.zeros(shape, np.uint8)
        nz, ny, nx = shape
        total = g.size
        solid = 0
        target = (1.0 - phi) * total
        while solid < target:
            c = rng.uniform(0, 1, 3) * np.array([nx, ny, nz])
            r = rng.uniform(rmin, rmax)
            lo = np.maximum(np.floor(c - r).astype(int), 0)
            hi = np.minimum(np.ceil(c + r).astype(int) + 1, [nx, ny, nz])
            if np.any(hi <= lo):
                continue
            zz, yy, xx = np.ogrid[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]]
            ball = (xx - c[0]) ** 2 + (yy - c[1]) ** 2 + (zz - c[2]) ** 2 <= r * r
            sub = g[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]]
            solid += int(np.count_nonzero(ball & (sub == 0)))
            sub[ball] = 1
        if not require_percolation or all(percolates(g, a) for a in axes):
            return g
    raise RuntimeError("no percolating sphere pack found")


def gaussian_field(n, corr_len: float, phi: float, seed: int,
                   require_percolation: bool = True, max_tries: int = 20, axes=(0,)) -> np.ndarray:
    """Thresholded Gaussian random field, fluid fraction phi; re-seeds until
    the fluid percolates along every array axis in ``axes`` (0 = z)."""
    from scipy import fft as sfft

    shape = (n, n, n) if np.isscalar(n) else tuple(n)
    for attempt in range(max_tries):
        rng = np.random.Generator(np.random.PCG64(seed + 1000 * attempt))
        noise = rng.standard_normal(shape, dtype=np.float32)
        spec = sfft.rfftn(noise, workers=-1)
        del noise
        kz = np.fft.fftfreq(shape[0]).astype(np.float32)[:, None, None] * np.float32(2 * np.pi)
        ky = np.fft.fftfreq(shape[1]).astype(np.float32)[None, :, None] * np.float32(2 * np.pi)
        kx = np.fft.rfftfreq(shape[2]).astype(np.float32)[None, None, :] * np.float32(2 * np.pi)
        l2 = np.float32(corr_len * corr_len / 8.0)
        spec *= np.exp(-(kz * kz) * l2)
        spec *= np.exp(-(ky * ky) * l2)
        spec *= np.exp(-(kx * kx) * l2)
        field = sfft.irfftn(spec, s=shape, workers=-1).astype(np.float32)
        del spec
        thr = np.quantile(field.ravel()[:: max(1, field.size // 4_000_000)], phi)
        g = np.where(field < thr, 0, 1).astype(np.uint8)
        del field
        if not require_percolation or all(percolates(g, a) for a in axes):
            return g
    raise RuntimeError("no percolating Gaussian field found")


def config0() -> np.ndarray:
    """54 x 54 x 50 box, solid except a straight pore along z centred at
    (27, 27) with radius 20 voxels."""
    return tube(54, 54, 50, 27, 27, 20)


def hagen_poiseuille_k(g: np.ndarray, r: float, dx: float, nz: int) -> float:
    """Analytic whole-domain k33 the reference reports for a straight tube.

    The fixed point of the sweep solves lap(w) = Re * dp/dz in voxel units,
    i.e. Poiseuille flow with viscosity 1/Re, under the reference's driving
    gradient G = (nz-1)/nz (inlet halo pinned to nz+2, last interior plane to
    3: src/boundary_conditions.cc:57-78).  Mean velocity Re*G*R^2/8, and
    k33 = mean/Re * phi * dx^2 (src/evaluation.cc:199-201) = G*R^2/8*phi*dx^2.
    """
    phi = float((g == 0).mean())
    grad = (nz - 1) / nz
    return r * r * grad / 8.0 * phi * dx * dx


def duct_k(w: float, h: float, dx: float, nz: int, phi: float, terms: int = 400) -> float:
    """Analytic whole-domain k33 of a rectangular duct of w x h voxels
    (series solution of laminar duct flow, White & Majdalani), same
    conventions as hagen_poiseuille_k."""
    grad = (nz - 1) / nz
    aa, bb = max(w, h), min(w, h)
    s = 0.0
    for m in range(terms):
        nn = 2 * m + 1
        s += np.tanh(nn * np.pi * aa / (2 * bb)) / nn ** 5
    mean = bb * bb / 12.0 * (1.0 - 192.0 * bb / (np.pi ** 5 * aa) * s) * grad
    return mean * phi * dx * dx


def random_media(count: int, side_min: int, side_max: int, seed: int, axes=(0,)):
    """``count`` cubic media with sides log-uniform in [side_min, side_max]:
    even indices sphere packs (radii 3-10 voxels), odd indices Gaussian random
    fields (correlation length uniform 4-12 voxels), porosity uniform
    0.15-0.45, each percolating along every axis in ``axes``.  Returns a list
    of dicts (raw, kind, phi, side, param)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        side = int(round(np.exp(rng.uniform(np.log(side_min), np.log(side_max)))))
        phi = float(rng.uniform(0.15, 0.45))
        s = int(rng.integers(1 << 30))
        if i % 2 == 0:
            raw = sphere_pack(side, phi, 3.0, 10.0, seed=s, axes=axes)
            out.append(dict(raw=raw, kind="spheres", phi=phi, side=side, param=(3.0, 10.0)))
        else:
            cl = float(rng.uniform(4.0, 12.0))
            # a level cut of a Gaussian field stops percolating near 16 % fluid:
            # raise the fraction in small steps until a percolating draw exists
            for bump in range(8):
                try:
                    raw = gaussian_field(side, cl, min(phi + 0.02 * bump, 0.45), seed=s, axes=axes,
                                         max_tries=4)
                    phi = min(phi + 0.02 * bump, 0.45)
                    break
                except RuntimeError:
                    continue
            else:
                raise RuntimeError("no percolating Gaussian field found")
            out.append(dict(raw=raw, kind="grf", phi=phi, side=side, param=cl))
    return out
