"""Synthetic scan geometries and seeded phantoms used by the tests and the benchmark.

Geometries (16-float parameter vector + view angles, the reference op attributes):

* ``small_scan``   256x256 image, 1.0 mm pixels, SID 500 mm, SDD 1000 mm, 512 equiangular
  bins of 1e-3 rad (1 mm arc pitch at the detector), 360 views over 360 deg.
* ``clinical_scan`` 512x512 image, 0.7 mm pixels (FOV radius 179.2 mm), SID 541 mm,
  SDD 949 mm, 888 bins of 1.0239 mm / 949 mm rad, 984 views over 360 deg.
* ``uneven_scans`` a seeded list of per-slice geometries of different image, view and
  detector sizes over the clinical FOV and fan (the uneven-batch workload).

Phantoms are sums of uniform ellipses (values added, then clipped), sampled at the
pixel centres, image layout [nY][nX].
"""
from __future__ import annotations

import math

import numpy as np

from .compat import angle, make_parameter

__all__ = ["small_scan", "clinical_scan", "scan", "ellipse_phantom", "clinical_phantom",
           "tiny_scan", "uneven_scans", "uneven_phantoms", "two_basis_batch", "poisson_target",
           "MU_WATER_PER_G_CC", "MU_IODINE_PER_MG_CC"]


def scan(*, nx, rfov, sid, sdd, n_cols, dcol_rad, n_views, n_slices=1, dshift=0.0, dleft_rad=None,
         ny=None, angular_range=360.0):
    """(parameter[16], angles[nViews]) for an equiangular fan-beam scan."""
    ny = nx if ny is None else ny
    dcol_deg = dcol_rad * 180.0 / math.pi
    dleft_deg = (-0.5 * n_cols * dcol_rad if dleft_rad is None else dleft_rad) * 180.0 / math.pi
    p = make_parameter(sdd=sdd, sod=sid, n_views=n_views, n_cols=n_cols, nx=nx, ny=ny, rfov=rfov,
                       dcol_deg=dcol_deg, dleft_deg=dleft_deg, dshift=dshift, angular_range=angular_range,
                       n_slices=n_slices)
    return p, angle(n_views, 0.0)


def small_scan(n_slices=1):
    return scan(nx=256, rfov=128.0, sid=500.0, sdd=1000.0, n_cols=512, dcol_rad=1e-3, n_views=360,
                n_slices=n_slices)


def clinical_scan(n_slices=1):
    return scan(nx=512, rfov=179.2, sid=541.0, sdd=949.0, n_cols=888, dcol_rad=1.0239 / 949.0, n_views=984,
                n_slices=n_slices)


CLINICAL_FAN_RAD = 888 * 1.0239 / 949.0  # 54.894 deg


def uneven_scans(n_slices=256, seed=4, sides=(256, 384, 512, 768), views=(360, 1200), bins=(512, 1024),
                 rfov=179.2, sid=541.0, sdd=949.0):
    """Per-slice (parameter, angles) of an uneven batch: image side uniform in `sides`,
    views uniform in [views[0], views[1]], bins uniform in [bins[0], bins[1]]; the FOV
    (radius rfov, pixel size 2 rfov / side) and the fan (54.894 deg, dCol = fan / bins)
    are fixed so every slice covers the same field."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_slices):
        side = int(rng.choice(sides))
        nv = int(rng.integers(views[0], views[1] + 1))
        nc = int(rng.integers(bins[0], bins[1] + 1))
        out.append(scan(nx=side, rfov=rfov, sid=sid, sdd=sdd, n_cols=nc, dcol_rad=CLINICAL_FAN_RAD / nc,
                        n_views=nv))
    return out


def uneven_phantoms(scans, seed=4, rfov=179.2):
    """One seeded ellipse phantom per slice of `scans` (seed + slice index), float32."""
    return [ellipse_phantom(int(p[5]), int(p[6]), rfov, 8, seed + 1000 + k) for k, (p, _) in enumerate(scans)]


# Linear attenuation per unit concentration at a 70 keV effective energy (mm^-1), from
# the NIST XCOM mass attenuation coefficients (water 0.193 cm^2/g; iodine about
# 4.9 cm^2/g at 70 keV, above its 33.2 keV K-edge).
MU_WATER_PER_G_CC = 0.0193
MU_IODINE_PER_MG_CC = 4.9e-4


def two_basis_batch(n_slices=16, nx=512, rfov=179.2, seed=100):
    """Two-basis images: water density U[0, 1.1] g/cc and iodine U[0, 20] mg/cc, each a
    seeded ellipse phantom (water seeds seed..seed+n-1, iodine seeds seed+n..seed+2n-1).
    Returns (water [n,nx,nx], iodine [n,nx,nx], mu = c_w water + c_I iodine), float32."""
    water = np.stack([ellipse_phantom(nx, nx, rfov, 8, seed + s, vmax=1.1) for s in range(n_slices)])
    iodine = np.stack([ellipse_phantom(nx, nx, rfov, 8, seed + n_slices + s, vmax=20.0) for s in range(n_slices)])
    mu = (MU_WATER_PER_G_CC * water.astype(np.float64) + MU_IODINE_PER_MG_CC * iodine).astype(np.float32)
    return water, iodine, mu


def poisson_target(line_integrals, photons=1e5, seed=7):
    """Noisy line integrals -ln(max(Poisson(I0 exp(-p)), 1) / I0), float64."""
    rng = np.random.default_rng(seed)
    counts = rng.poisson(photons * np.exp(-np.asarray(line_integrals, dtype=np.float64)))
    return -np.log(np.maximum(counts, 1) / photons)


def tiny_scan(nx=16, n_views=24, n_cols=20, n_slices=1, sid=60.0, rfov=8.0, fan_margin=1.15, dshift=0.0):
    """A small geometry whose fan covers the FOV with margin (for brute-force checks)."""
    half_fan = math.asin(rfov / sid) * fan_margin
    dcol = 2.0 * half_fan / n_cols
    return scan(nx=nx, rfov=rfov, sid=sid, sdd=2.0 * sid, n_cols=n_cols, dcol_rad=dcol, n_views=n_views,
                n_slices=n_slices, dshift=dshift)


def ellipse_phantom(nx, ny, rfov, n_ellipses=8, seed=0, vmax=0.04):
    """Sum of seeded uniform ellipses: centres uniform in a disk of radius 0.7*rFOV,
    semi-axes U[0.05, 0.4]*rFOV, orientation U[0, pi), values U[0, vmax]; clipped to
    [0, vmax].  float32 [ny, nx]."""
    rng = np.random.default_rng(seed)
    dx = 2.0 * rfov / nx
    dy = 2.0 * rfov / ny
    x = -rfov + (np.arange(nx) + 0.5) * dx
    y = -rfov + (np.arange(ny) + 0.5) * dy
    X, Y = np.meshgrid(x, y)
    img = np.zeros((ny, nx), dtype=np.float64)
    for _ in range(n_ellipses):
        rr = 0.7 * rfov * math.sqrt(rng.uniform())
        th = rng.uniform(0, 2 * math.pi)
        cx, cy = rr * math.cos(th), rr * math.sin(th)
        a, b = rng.uniform(0.05, 0.4, size=2) * rfov
        phi = rng.uniform(0, math.pi)
        val = rng.uniform(0, vmax)
        c, s = math.cos(phi), math.sin(phi)
        xr = (X - cx) * c + (Y - cy) * s
        yr = -(X - cx) * s + (Y - cy) * c
        img[(xr / a) ** 2 + (yr / b) ** 2 <= 1.0] += val
    return np.clip(img, 0.0, vmax).astype(np.float32)


def clinical_phantom(n_slices=1, seed=1, noise_seed=2, noise_sigma=1e-4, nx=512, rfov=179.2):
    """Ellipse phantoms (seeds seed, seed+1, ...) plus N(0, noise_sigma) texture noise."""
    rng = np.random.default_rng(noise_seed)
    out = np.empty((n_slices, nx, nx), dtype=np.float32)
    for s in range(n_slices):
        out[s] = ellipse_phantom(nx, nx, rfov, 8, seed + s)
        out[s] += rng.normal(0.0, noise_sigma, size=(nx, nx)).astype(np.float32)
    return out
