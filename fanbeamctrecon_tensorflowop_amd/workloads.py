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
           "MU_WATER_PER_G_CC", "MU_IODINE_PER_MG_CC", "cone_scan", "random_ellipsoids",
           "ellipsoid_volume", "ellipsoid_projections"]


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


def cone_scan(*, nx=512, nz=256, dx=0.7, dz=None, n_views=720, n_cols=1024, n_rows=256, pitch=1.0,
              row_pitch=None, R=541.0, D=949.0, detector="flat", short_scan=False):
    """(geometry dict of fbct_cone_geom, angles) for a circular orbit: full 360 degree
    scan of n_views views, a centred detector of n_cols x n_rows cells (pitch mm for the
    flat detector at distance D; the curved one spans the same fan), and a centred
    nx x nx x nz volume of dx x dx x dz voxels (FOV radius nx dx / 2).  The default is
    the 512 x 512 x 256 volume, 720 views, 1024 x 256 flat panel at 1.0 mm workload.
    short_scan: the views span 180 degrees plus the fan angle (plus one view spacing) at
    the full-scan spacing of n_views views."""
    from .cone import cone_geom
    dz = dx if dz is None else dz
    row_pitch = pitch if row_pitch is None else row_pitch
    if detector == "flat":
        d_col = pitch
        half_fan = math.atan(0.5 * n_cols * pitch / D)
    else:
        d_col = 2.0 * math.atan(0.5 * n_cols * pitch / D) / n_cols
        half_fan = 0.5 * n_cols * d_col
    step = 2.0 * math.pi / n_views
    angles = angle(n_views, 0.0)
    if short_scan:
        angles = angles[:int(math.ceil((math.pi + 2.0 * half_fan) / step)) + 1]
    g = cone_geom(R=R, D=D, n_views=len(angles), d_angle=step, detector=detector, n_cols=n_cols,
                  n_rows=n_rows, d_col=d_col, d_row=row_pitch, nx=nx, nz=nz, dx=dx, dz=dz, short_scan=short_scan)
    return g, angles


def random_ellipsoids(n=10, seed=5, rfov=179.2, zhalf=89.6, vmax=0.04):
    """Seeded ellipsoids [n, 8] = (cx, cy, cz, a, b, c, phi, value): centres uniform in a
    cylinder of radius 0.6 rfov and half-height 0.6 zhalf, semi-axes U[0.05, 0.35] rfov
    (c: U[0.1, 0.6] zhalf), rotation phi about z uniform in [0, pi), value U[0, vmax]."""
    rng = np.random.default_rng(seed)
    out = np.zeros((n, 8))
    for k in range(n):
        rr = 0.6 * rfov * math.sqrt(rng.uniform())
        th = rng.uniform(0, 2 * math.pi)
        out[k, 0:3] = rr * math.cos(th), rr * math.sin(th), rng.uniform(-0.6, 0.6) * zhalf
        out[k, 3:5] = rng.uniform(0.05, 0.35, size=2) * rfov
        out[k, 5] = rng.uniform(0.1, 0.6) * zhalf
        out[k, 6] = rng.uniform(0, math.pi)
        out[k, 7] = rng.uniform(0, vmax)
    return out


def ellipsoid_volume(ells, geom):
    """Sum of the ellipsoids' values at the voxel centres, float32 [nZ, nY, nX]."""
    g = geom
    x = g["xmin"] + (np.arange(g["nX"]) + 0.5) * g["dx"]
    y = g["ymin"] + (np.arange(g["nY"]) + 0.5) * g["dy"]
    X, Y = np.meshgrid(x, y)
    vol = np.zeros((g["nZ"], g["nY"], g["nX"]), dtype=np.float32)
    for k in range(g["nZ"]):
        z = g["zmin"] + (k + 0.5) * g["dz"]
        sl = np.zeros((g["nY"], g["nX"]))
        for cx, cy, cz, a, b, c, phi, val in ells:
            if abs(z - cz) >= c:
                continue
            cp, sp = math.cos(phi), math.sin(phi)
            xr = (X - cx) * cp + (Y - cy) * sp
            yr = -(X - cx) * sp + (Y - cy) * cp
            sl[(xr / a) ** 2 + (yr / b) ** 2 + ((z - cz) / c) ** 2 <= 1.0] += val
        vol[k] = sl
    return vol


def ellipsoid_projections(ells, geom, angles, device="cpu"):
    """Exact line integrals of the ellipsoids for every detector cell centre, float64
    [nViews, nRows, nCols] (a torch tensor on `device`).

    Source at R (cos b, sin b, 0); the cell (u, v) lies at S + D e_c + u e_u + v e_z with
    e_c = -(cos b, sin b, 0), e_u = (sin b, -cos b, 0): the flat panel of the cone
    geometry; for the curved detector u = D tan(gamma) of the column's fan angle."""
    import torch
    g = geom
    dt = torch.float64
    nC, nR = g["nCols"], g["nRows"]
    cpos = g["colLeft"] + (torch.arange(nC, dtype=dt, device=device) + 0.5 - g["dShift"]) * g["dCol"]
    u = cpos if g["detector"] == 0 else g["D"] * torch.tan(cpos)
    v = g["rowBottom"] + (torch.arange(nR, dtype=dt, device=device) + 0.5) * g["dRow"]
    V, Uc = torch.meshgrid(v, u, indexing="ij")
    out = torch.zeros((len(angles), nR, nC), dtype=dt, device=device)
    R, D = float(g["R"]), float(g["D"])
    for vi, b in enumerate(np.asarray(angles, dtype=np.float64)):
        cb, sb = math.cos(float(np.float32(b))), math.sin(float(np.float32(b)))
        S = torch.tensor([R * cb, R * sb, 0.0], dtype=dt, device=device)
        # direction to the cell: D e_c + u e_u + v e_z
        dx_ = -D * cb + Uc * sb
        dy_ = -D * sb - Uc * cb
        dz_ = V
        nrm = torch.sqrt(dx_ * dx_ + dy_ * dy_ + dz_ * dz_)
        dx_, dy_, dz_ = dx_ / nrm, dy_ / nrm, dz_ / nrm
        acc = torch.zeros_like(nrm)
        for cx, cy, cz, a, bb, c, phi, val in ells:
            cp, sp = math.cos(phi), math.sin(phi)
            # ellipsoid frame: rotate by -phi about z, scale by the semi-axes
            ox, oy, oz = float(S[0]) - cx, float(S[1]) - cy, -cz
            o1, o2, o3 = (ox * cp + oy * sp) / a, (-ox * sp + oy * cp) / bb, oz / c
            d1 = (dx_ * cp + dy_ * sp) / a
            d2 = (-dx_ * sp + dy_ * cp) / bb
            d3 = dz_ / c
            A = d1 * d1 + d2 * d2 + d3 * d3
            B = 2.0 * (d1 * o1 + d2 * o2 + d3 * o3)
            C = o1 * o1 + o2 * o2 + o3 * o3 - 1.0
            disc = B * B - 4.0 * A * C
            acc += val * torch.where(disc > 0, torch.sqrt(torch.clamp(disc, min=0.0)) / A, torch.zeros_like(A))
        out[vi] = acc
    return out


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
