"""Synthetic PET workloads: scanner, sinogram and image geometries plus
seeded ellipsoid phantoms in a water cylinder.

The geometries are the benchmark configurations of this project:

* ``c1``: 4 rings x 128 crystals, radius 150 mm, 4 mm axial pitch, span 1,
  7 segments, 65 tangential bins (crystal separation >= 32); 64x64x8 image
  at 2 mm.  66 560 LORs.
* ``c2``: 24 rings x 448 crystals, radius 330 mm, 4 mm pitch, span 1,
  47 segments (all ring differences), 325 tangential bins (600 mm
  transaxial FOV); 160x160x47 image at 2x2x3.27 mm.  41 932 800 LORs.
* ``c5``: 64 rings x 672 crystals, radius 410 mm, 3.2 mm pitch, 127
  segments, 225 tangential bins; 256x256x128 image at 1.6 mm.

Phantoms: ellipsoids with uniform activity inside a water cylinder whose
Hounsfield map is 0 inside and -1000 outside (water mu = 0.0096/mm after
the HU conversion of the reference, operations.cc:243-284).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .interfile import ProjSpec, ScannerSpec, VolSpec


@dataclass
class Workload:
    name: str
    scanner: ScannerSpec
    proj: ProjSpec
    vol: VolSpec
    water_radius: float
    n_ellipsoids: int
    activity: tuple
    centre_r: float
    centre_z: float
    semi_xy: tuple
    semi_z: tuple
    counts: float
    seed: int
    n_iterations: int
    n_subsets: int
    fwhm: tuple = (4.0, 4.0, 4.0)

    @property
    def n_lors(self) -> int:
        return self.proj.geometry()[3]


def c1() -> Workload:
    return Workload(
        name="c1_4x128_64x64x8",
        scanner=ScannerSpec(rsector_repeat_number=128,
                            rsector_inner_radius=150.0,
                            crystal_dimensions=(0.0, 0.0, 4.0),
                            crystal_repeat_numbers=(1, 4)),
        proj=ProjSpec(n_rings=4, n_crystals_per_ring=128, n_segments=7,
                      n_tang_coords=65),
        vol=VolSpec(size=(64, 64, 8), voxel=(2.0, 2.0, 2.0),
                    offset=(-63.0, -63.0, 0.0)),
        water_radius=60.0, n_ellipsoids=3, activity=(0.5, 4.0),
        centre_r=40.0, centre_z=6.0, semi_xy=(5.0, 15.0), semi_z=(2.0, 6.0),
        counts=1e5, seed=0, n_iterations=2, n_subsets=4)


def c2() -> Workload:
    return Workload(
        name="c2_24x448_160x160x47",
        scanner=ScannerSpec(rsector_repeat_number=448,
                            rsector_inner_radius=330.0,
                            crystal_dimensions=(0.0, 0.0, 4.0),
                            crystal_repeat_numbers=(1, 24)),
        proj=ProjSpec(n_rings=24, n_crystals_per_ring=448, n_segments=47,
                      n_tang_coords=325),
        vol=VolSpec(size=(160, 160, 47), voxel=(2.0, 2.0, 3.27),
                    offset=(-159.0, -159.0, 0.0)),
        water_radius=100.0, n_ellipsoids=12, activity=(0.1, 10.0),
        centre_r=70.0, centre_z=40.0, semi_xy=(8.0, 30.0),
        semi_z=(5.0, 20.0), counts=5e8, seed=1, n_iterations=3,
        n_subsets=16)


def c5() -> Workload:
    return Workload(
        name="c5_64x672_256x256x128",
        scanner=ScannerSpec(rsector_repeat_number=672,
                            rsector_inner_radius=410.0,
                            crystal_dimensions=(0.0, 0.0, 3.2),
                            crystal_repeat_numbers=(1, 64)),
        proj=ProjSpec(n_rings=64, n_crystals_per_ring=672, n_segments=127,
                      n_tang_coords=225),
        vol=VolSpec(size=(256, 256, 128), voxel=(1.6, 1.6, 1.6),
                    offset=(-204.0, -204.0, 0.0)),
        water_radius=150.0, n_ellipsoids=16, activity=(0.1, 10.0),
        centre_r=100.0, centre_z=60.0, semi_xy=(8.0, 35.0),
        semi_z=(5.0, 25.0), counts=1e9, seed=5, n_iterations=2,
        n_subsets=24)


WORKLOADS = {"c1": c1, "c2": c2, "c5": c5}


def voxel_centres(vol: VolSpec):
    """Voxel-centre coordinates (x, y, z) in mm, each 1-D."""
    nx, ny, nz = vol.size
    dx, dy, dz = vol.voxel
    x = vol.offset[0] + dx * np.arange(nx)
    y = vol.offset[1] + dy * np.arange(ny)
    z = vol.offset[2] - nz * dz / 2.0 + dz * (np.arange(nz) + 0.5)
    return x, y, z


def phantom(w: Workload, seed: int | None = None):
    """Seeded activity image and HU map, each (nz, ny, nx) float32."""
    rng = np.random.default_rng(w.seed if seed is None else seed)
    x, y, z = voxel_centres(w.vol)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    act = np.zeros(Z.shape, dtype=np.float64)
    for _ in range(w.n_ellipsoids):
        r = w.centre_r * np.sqrt(rng.uniform())
        th = rng.uniform(0, 2 * np.pi)
        cx, cy = r * np.cos(th), r * np.sin(th)
        cz = rng.uniform(-w.centre_z, w.centre_z)
        ax, ay = rng.uniform(*w.semi_xy, size=2)
        az = rng.uniform(*w.semi_z)
        a = rng.uniform(*w.activity)
        inside = (((X - cx) / ax) ** 2 + ((Y - cy) / ay) ** 2
                  + ((Z - cz) / az) ** 2) <= 1.0
        act[inside] = a
    water = (X ** 2 + Y ** 2) <= w.water_radius ** 2
    act[~water] = 0.0
    hu = np.where(water, 0.0, -1000.0)
    return act.astype(np.float32), hu.astype(np.float32)


def poisson_counts(y0: np.ndarray, acf: np.ndarray, total: float, seed: int):
    """y ~ Poisson(total * a*y0 / sum(a*y0)) with a = 1/ACF (attenuated
    true counts), as float32."""
    rng = np.random.default_rng(seed)
    a = np.where(acf > 0, 1.0 / np.maximum(acf, 1e-30), 0.0)
    mean = a * y0.astype(np.float64)
    s = mean.sum()
    mean = mean * (total / s) if s > 0 else mean
    return rng.poisson(mean).astype(np.float32)


def attenuation(mu_line: np.ndarray) -> np.ndarray:
    """exp(-m) of line integrals of mu, in float64: the reciprocal of
    exp(m) from its Taylor polynomial of degree 30 (positive terms only),
    evaluated with plain IEEE multiplies, divides and adds (Horner).

    The data of the parity runs must come out bit for bit the same on every
    machine, and vectorised libm exp may differ in the last place between
    CPUs; these operations do not. For 0 <= m <= 4 the result is within
    1e-15 relative of exp(-m)."""
    m = np.asarray(mu_line, np.float64)
    acc = np.ones_like(m)
    for k in range(30, 0, -1):
        acc = acc * m
        acc = acc / float(k)
        acc = acc + 1.0
    return 1.0 / acc


def parity_counts(w: Workload, y0: np.ndarray, mu_line: np.ndarray):
    """Attenuated Poisson counts of a workload with bit-reproducible inputs:
    y ~ Poisson(scale * y0 * exp(-m)), seeded with ``w.seed``, where y0 is
    the forward projection of the phantom, m that of its mu map, and scale
    makes the expected total ``w.counts`` (rounded to 6 significant digits,
    so that summation order cannot change it). Returns (counts float32,
    scale)."""
    lam = y0.astype(np.float64) * attenuation(mu_line)
    s = float(lam.sum())
    scale = float(f"{w.counts / s:.6g}") if s > 0 else 0.0
    lam *= scale
    rng = np.random.default_rng(w.seed)
    return rng.poisson(lam).astype(np.float32), scale


def hash_counts(n_bins: int) -> np.ndarray:
    """Deterministic count sinogram for the precision runs of scanners
    whose data is too large to ship: y[b] = (b * 2654435761 mod 2^32) >> 27,
    integers 0-31, identical on every host."""
    y = np.empty(n_bins, np.float32)
    step = 1 << 26
    for a in range(0, n_bins, step):
        b = np.arange(a, min(a + step, n_bins), dtype=np.uint64)
        y[a:a + b.size] = (((b * np.uint64(2654435761)) & np.uint64(0xffffffff))
                           >> np.uint64(27)).astype(np.float32)
    return y


def image_summary(img: np.ndarray, shape, block: int = 8) -> dict:
    """Compact view of an image for fixtures of large grids: sums over
    block^3 voxel blocks and the three central planes (z, y and x through
    the grid centre), as float64. shape = (nz, ny, nx)."""
    a = np.asarray(img, np.float64).reshape(shape)
    nz, ny, nx = shape
    bz, by, bx = nz // block, ny // block, nx // block
    blocks = a[:bz * block, :by * block, :bx * block].reshape(
        bz, block, by, block, bx, block).sum(axis=(1, 3, 5))
    return {"blocks": blocks, "plane_z": a[nz // 2].copy(),
            "plane_y": a[:, ny // 2, :].copy(),
            "plane_x": a[:, :, nx // 2].copy()}
