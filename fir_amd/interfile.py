"""Key-file and Interfile readers/writers for FIR's on-disk formats.

Python-side mirror of the formats the FIR executables read and write:

* scanner ``.scan`` files (keys of reference ScannerInterfileReader.cc:5-52),
* sinogram ``.hs`` headers plus raw native FP32 ``.s`` data in segment order
  (reference ProjInterfileReader.cc:13-47,137-192),
* volume ``.h33`` headers plus ``.i33`` data, written as FLOAT / 4 bytes /
  LITTLEENDIAN with ``%.7f`` doubles (reference VolInterfileReader.cc:551-661,
  writeKeys.cc:12-21),
* parameter files for FIR_ForwardProj / FIR_OSEM (reference
  ForwardProj.cc:111-149, OSEM.cc:385-450).

The text layout follows the reference's PyInteractor writers
(interfile_io.py:62-84, vol.py:83-138, proj.py:69-89, scanner.py:109-157):
``key := value`` lines between ``!TYPE :=`` and ``!END OF TYPE :=``.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np


def _fmt_list(values) -> str:
    return "{" + ",".join(str(v) for v in values) + "}"


def write_keys(path: str, file_type: str, lines: Sequence[Optional[tuple]]):
    """Write an Interfile-style key file (one ``key := value`` per line)."""
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    out = [f"!{file_type} :=\n"]
    for el in lines:
        out.append("\n" if el is None else f"{el[0]} := {el[1]}\n")
    out.append(f"!END OF {file_type} :=\n")
    with open(path, "w") as f:
        f.writelines(out)


@dataclass
class ScannerSpec:
    """Cylindrical scanner keys (reference ScannerHeader.h:7-46)."""

    rsector_repeat_number: int
    rsector_inner_radius: float
    crystal_dimensions: Sequence[float] = (0.0, 0.0, 0.0)
    crystal_repeat_numbers: Sequence[int] = (1, 1)
    inter_crystal_distance: Sequence[float] = (0.0, 0.0)
    module_dimensions: Sequence[float] = (0.0, 0.0, 0.0)
    module_repeat_numbers: Sequence[int] = (1, 1)
    inter_module_distance: Sequence[float] = (0.0, 0.0)
    rsector_dimensions: Sequence[float] = (0.0, 0.0, 0.0)

    @property
    def n_rings(self) -> int:
        return self.crystal_repeat_numbers[1] * self.module_repeat_numbers[1]

    @property
    def n_crystals_per_ring(self) -> int:
        return (self.crystal_repeat_numbers[0] * self.module_repeat_numbers[0]
                * self.rsector_repeat_number)

    def header_vector(self) -> np.ndarray:
        """The 19 doubles the geometry routines take (same order as keys)."""
        v = (list(self.crystal_dimensions) + list(self.crystal_repeat_numbers)
             + list(self.inter_crystal_distance) + list(self.module_dimensions)
             + list(self.module_repeat_numbers)
             + list(self.inter_module_distance)
             + list(self.rsector_dimensions)
             + [self.rsector_repeat_number, self.rsector_inner_radius])
        return np.asarray(v, dtype=np.float64)

    def write(self, path: str):
        write_keys(path, "SCANNER PARAMETERS", [
            ("crystal dimensions XYZ in mm", _fmt_list(self.crystal_dimensions)),
            ("crystal repeat numbers YZ", _fmt_list(self.crystal_repeat_numbers)),
            ("inter-crystal distance YZ in mm",
             _fmt_list(self.inter_crystal_distance)),
            None,
            ("module dimensions XYZ in mm", _fmt_list(self.module_dimensions)),
            ("module repeat numbers YZ", _fmt_list(self.module_repeat_numbers)),
            ("inter-module distance YZ in mm",
             _fmt_list(self.inter_module_distance)),
            None,
            ("rSector dimensions XYZ in mm", _fmt_list(self.rsector_dimensions)),
            ("rSector repeat number", str(self.rsector_repeat_number)),
            ("rSector inner radius in mm", str(self.rsector_inner_radius)),
        ])


@dataclass
class ProjSpec:
    """Sinogram header keys (reference ProjHeader.h:15-40)."""

    n_rings: int
    n_crystals_per_ring: int
    n_segments: int = 1
    n_tang_coords: int = 0
    segment_span: int = 1

    def write(self, path: str, data_file: Optional[str] = None):
        lines = []
        if data_file is not None:
            lines.append(("name of data file", data_file))
        lines += [
            ("number of rings", str(self.n_rings)),
            ("number of crystals per ring", str(self.n_crystals_per_ring)),
            ("segment span", str(self.segment_span)),
            ("number of segments", str(self.n_segments)),
            ("number of tangential coordinates", str(self.n_tang_coords)),
        ]
        write_keys(path, "PROJECTION DATA PARAMETERS", lines)

    def geometry(self):
        """(n_tang, n_views, n_axial list, n_bins) as ProjGeometry::fill."""
        nt = self.n_tang_coords or self.n_crystals_per_ring - 1
        so = (self.n_segments - 1) // 2
        if self.segment_span == 1:
            na = [self.n_rings - abs(s) for s in range(-so, so + 1)]
        else:
            na = []
            for s in range(-so, so + 1):
                n = 2 * self.n_rings - 1
                if abs(s) >= 1:
                    n -= self.segment_span + 1
                if abs(s) >= 2:
                    n -= 2 * self.segment_span * (abs(s) - 1)
                na.append(n)
        nv = self.n_crystals_per_ring // 2
        return nt, nv, na, sum(na) * nv * nt


@dataclass
class VolSpec:
    """Volume header keys (reference VolHeader.h:6-34)."""

    size: Sequence[int]
    voxel: Sequence[float]
    offset: Sequence[float]
    n_frames: int = 1

    @property
    def n_voxels(self) -> int:
        return int(self.size[0]) * int(self.size[1]) * int(self.size[2])

    def write(self, header_path: str, data: Optional[np.ndarray] = None):
        """Write ``.h33`` (and ``.i33`` when data is given)."""
        lines = []
        if data is not None:
            base = os.path.splitext(header_path)[0]
            data_path = base + ".i33"
            lines.append(("name of data file", os.path.basename(data_path)))
            np.ascontiguousarray(data, dtype="<f4").tofile(data_path)
        lines += [
            ("data offset in bytes", "0"),
            ("number format", "FLOAT"),
            ("number of bytes per pixel", "4"),
            ("imagedata byte order", "LITTLEENDIAN"),
            None,
            ("number of dimensions", "3"),
            None,
            ("matrix size [1]", str(self.size[0])),
            ("matrix size [2]", str(self.size[1])),
            ("matrix size [3]", str(self.size[2])),
            ("number of slices", str(self.size[2])),
            None,
            ("scaling factor (mm/pixel) [1]", f"{self.voxel[0]:.7f}"),
            ("scaling factor (mm/pixel) [2]", f"{self.voxel[1]:.7f}"),
            ("scaling factor (mm/pixel) [3]", f"{self.voxel[2]:.7f}"),
            ("slice thickness (pixels)", f"{self.voxel[2]:.7f}"),
            None,
            ("first pixel offset (mm) [1]", f"{self.offset[0]:.7f}"),
            ("first pixel offset (mm) [2]", f"{self.offset[1]:.7f}"),
            ("first pixel offset (mm) [3]", f"{self.offset[2]:.7f}"),
            None,
            ("number of time frames", str(self.n_frames)),
        ]
        write_keys(header_path, "INTERFILE", lines)


def read_keys(path: str) -> dict:
    """Parse ``key := value`` lines into a dict with standardised keys
    (lower case, runs of space/tab/underscore/! collapsed), as the
    reference KeyParser::standardiseKeyword does (KeyParser.cc:589-651)."""
    out = {}
    with open(path) as f:
        for line in f:
            if ":" not in line:
                continue
            key = line.split(":", 1)[0]
            key = " ".join(key.replace("_", " ").replace("!", " ").split())
            key = key.lower()
            value = line.split("=", 1)[1].strip() if "=" in line else ""
            out[key] = value
    return out


def read_vol(header_path: str) -> np.ndarray:
    """Read a FLOAT/4 little-endian volume as (frames, nz, ny, nx)."""
    keys = read_keys(header_path)
    nx = int(keys["matrix size [1]"])
    ny = int(keys["matrix size [2]"])
    nz = int(keys["number of slices"])
    nf = int(keys.get("number of time frames", "1") or 1)
    data = keys["name of data file"]
    if not data.startswith("/"):
        data = os.path.join(os.path.dirname(header_path), data)
    arr = np.fromfile(data, dtype="<f4", count=nx * ny * nz * nf)
    return arr.reshape(nf, nz, ny, nx)


def read_proj(header_path: str) -> np.ndarray:
    """Read the raw FP32 sinogram (segment-concatenated, flat)."""
    keys = read_keys(header_path)
    data = keys["name of data file"]
    if not data.startswith("/"):
        data = os.path.join(os.path.dirname(header_path), data)
    return np.fromfile(data, dtype=np.float32)


def write_proj(header_path: str, spec: ProjSpec, data: np.ndarray):
    base = os.path.splitext(header_path)[0]
    spec.write(header_path, os.path.basename(base + ".s"))
    np.ascontiguousarray(data, dtype=np.float32).tofile(base + ".s")


def write_params(path: str, start: str, items: Sequence[tuple]):
    """Parameter file for FIR_ForwardProj ("FORWARD PROJECTION PARAMETERS")
    or FIR_OSEM ("OSEM PARAMETERS")."""
    write_keys(path, start, [(k, v) for k, v in items])
