"""The C++ host layer (fir_amd/libfir_host.so) without a GPU: file formats
byte-compatible with the reference writers, the key-file grammar, scanner
tables, sinogram geometry and the LOR-cache order (fir_host_selftest)."""
import json
import os
import subprocess

import numpy as np
import pytest

import pyoracle as O
from conftest import GOLDEN, ROOT
from fir_amd import interfile as io
from fir_amd import synth

SELFTEST = os.path.join(ROOT, "fir_amd", "bin", "fir_host_selftest")


def run(*args, cwd=None, check=True, env=None):
    if not os.path.exists(SELFTEST):
        pytest.fail("host layer not built (python -m fir_amd.build)")
    r = subprocess.run([SELFTEST, *map(str, args)], capture_output=True,
                       text=True, cwd=cwd, env=env)
    if check and r.returncode != 0:
        raise AssertionError(r.stderr)
    return r


@pytest.fixture(scope="module")
def headers():
    return json.load(open(os.path.join(GOLDEN, "ref_headers.json")))


def test_volume_writer_matches_reference_bytes(tmp_path, headers, c1_golden):
    w = synth.c1()
    w.vol.write(str(tmp_path / "in.h33"), c1_golden["bp"])
    run("vol", tmp_path / "in.h33", tmp_path / "bp", cwd=tmp_path)
    assert (tmp_path / "bp.h33").read_text() == headers["bp.h33"]
    data = np.fromfile(tmp_path / "bp.i33", "<f4")
    assert np.array_equal(data, c1_golden["bp"].ravel())


def test_multiframe_volume_roundtrip(tmp_path, headers, c1_golden):
    w = synth.c1()
    spec = io.VolSpec(w.vol.size, w.vol.voxel, w.vol.offset, n_frames=4)
    spec.write(str(tmp_path / "s.h33"), c1_golden["sens"])
    run("vol", tmp_path / "s.h33", tmp_path / "sens", cwd=tmp_path)
    assert (tmp_path / "sens.h33").read_text() == headers["sens.h33"]
    assert np.array_equal(np.fromfile(tmp_path / "sens.i33", "<f4"),
                          c1_golden["sens"].ravel())


def test_sinogram_writer_matches_reference_bytes(tmp_path, headers,
                                                 c1_golden):
    w = synth.c1()
    io.write_proj(str(tmp_path / "in.hs"), w.proj, c1_golden["y0"])
    run("proj", tmp_path / "in.hs", tmp_path / "y0", cwd=tmp_path)
    assert (tmp_path / "y0.hs").read_text() == headers["y0.hs"]
    assert np.array_equal(np.fromfile(tmp_path / "y0.s", "<f4"),
                          c1_golden["y0"])


@pytest.mark.parametrize("fmt,bpp,order,dtype", [
    ("unsigned integer", 1, "LITTLEENDIAN", "<u1"),
    ("unsigned integer", 2, "BIGENDIAN", ">u2"),
    ("signed integer", 4, "bigendian", ">i4"),
    ("signed integer", 8, "LITTLEENDIAN", "<i8"),
    ("long float", 8, "BIGENDIAN", ">f8"),
    ("short float", 4, "LITTLEENDIAN", "<f4"),
])
def test_volume_reader_types(tmp_path, fmt, bpp, order, dtype):
    vals = (np.arange(2 * 3 * 4) + (1 if "u" in dtype else -3)).astype(dtype)
    (tmp_path / "d.raw").write_bytes(b"\0" * 16 + vals.tobytes())
    io.write_keys(str(tmp_path / "a.h33"), "INTERFILE", [
        ("name of data file", "d.raw"), ("data offset in bytes", "16"),
        ("number format", fmt), ("number of bytes per pixel", str(bpp)),
        ("imagedata byte order", order), ("matrix size [1]", "2"),
        ("matrix size [2]", "3"), ("number of slices", "4"),
        ("scaling factor (mm/pixel) [1]", "1"),
        ("scaling factor (mm/pixel) [2]", "1"),
        ("slice thickness (pixels)", "1")])
    run("vol", tmp_path / "a.h33", tmp_path / "o", cwd=tmp_path)
    got = np.fromfile(tmp_path / "o.i33", "<f4")
    assert np.array_equal(got, vals.astype(np.float32))


def test_short_data_file_is_an_error(tmp_path):
    w = synth.c1()
    w.vol.write(str(tmp_path / "in.h33"), np.zeros(10, np.float32))
    r = run("vol", tmp_path / "in.h33", tmp_path / "o", check=False)
    assert r.returncode != 0 and "inferior" in r.stderr


def test_scanner_tables_match_reference(tmp_path, tables_golden):
    for name, sc in (("c1", synth.c1().scanner), ("c2", synth.c2().scanner)):
        sc.write(str(tmp_path / f"{name}.scan"))
        out = run("scanner", tmp_path / f"{name}.scan").stdout.split("\n")
        n, ns = map(int, out[0].split())
        xy = np.array([float(v) for line in out[1:1 + n]
                       for v in line.split()])
        z = np.array([float(v) for v in out[1 + n:1 + n + ns]])
        assert np.array_equal(xy, tables_golden[f"{name}_xy"])
        assert np.array_equal(z, tables_golden[f"{name}_z"])


@pytest.mark.parametrize("args,expect", [
    ((5, 32, 1, 3, 8), "1664 1 4 16 0 1 8 4 5 4"),
    ((5, 32, 3, 3, 8), "2432 1 4 16 1 4 8 5 9 5"),
    ((4, 128, 1, 7, 0), "130048 3 63 64 0 3 127 1 2 3 4 3 2 1"),
])
def test_proj_geometry_known_answers(args, expect):
    assert run("geom", *args).stdout.strip() == expect


@pytest.mark.parametrize("args,msg", [
    ((0, 4, 1, 1, 0), "number of rings"),
    ((1, 6, 1, 1, 0), "multiple of four"),
    ((2, 8, 2, 1, 0), "odd number"),
    ((2, 8, 1, 5, 0), "must not be > 3"),
    ((2, 8, 1, 1, 8), "must not be > 7"),
])
def test_proj_geometry_rejects(args, msg):
    r = run("geom", *args, check=False)
    assert r.returncode != 0 and msg in r.stderr


def test_lor_cache_order(tmp_path):
    """The cache enumerates (subview, axial, tang) per (subset, segment)
    with view = subview * S + subset; crystals match the oracle mapping."""
    w = synth.c1()
    w.proj.write(str(tmp_path / "p.hs"))
    out = run("lors", tmp_path / "p.hs", 4).stdout.split("\n")[:-1]
    rows = np.array([list(map(int, l.split())) for l in out])
    assert len(rows) == w.n_lors
    p = O.proj_desc(w.proj)
    k = 0
    for s in range(4):
        for seg in range(-3, 4):
            na = p.n_axial[seg + 3]
            for sv in range(16):
                view = sv * 4 + s
                for ax in range(na):
                    for t in range(-32, 33):
                        r = rows[k]
                        k += 1
                        if (sv * na + ax) % 7 or t % 5:
                            continue
                        s1, s2 = O.crystal_axial(p, seg, ax)
                        a1, a2 = O.crystal_ang(p, view, t)
                        assert list(r) == [seg, 1,
                                           (view * na + ax) * 65 + t + 32,
                                           s1, a1, s2, a2]
    assert k == len(rows)


@pytest.mark.parametrize("geom", [
    dict(n_rings=4, n_crystals_per_ring=128, n_segments=7, n_tang_coords=65),
    dict(n_rings=8, n_crystals_per_ring=64, n_segments=5, n_tang_coords=49,
         segment_span=3),
    dict(n_rings=8, n_crystals_per_ring=32, n_segments=3, n_tang_coords=15,
         segment_span=5)])
def test_bin_coordinates_match_oracle(tmp_path, geom):
    """ProjData::getBinCoordinates of the host layer equals the oracle's
    (itself pinned to the compiled reference) for every crystal pair."""
    spec = io.ProjSpec(**geom)
    spec.write(str(tmp_path / "p.hs"))
    run("bins", tmp_path / "p.hs", tmp_path / "b.bin")
    got = np.fromfile(tmp_path / "b.bin", np.int32).reshape(-1, 5)
    p = O.proj_desc(spec)
    nr, nc = spec.n_rings, spec.n_crystals_per_ring
    assert got.shape[0] == (nr * nc) ** 2
    k = 0
    for r1 in range(nr):
        for c1 in range(nc):
            for r2 in range(nr):
                for c2 in range(nc):
                    b = O.bin_coords(p, r1, c1, r2, c2)
                    g = got[k]
                    k += 1
                    assert (b is None) == (g[0] == 0)
                    if b is not None:
                        assert tuple(g[1:]) == b


@pytest.mark.parametrize("raw,std", [
    ("!Number_of   Rings", "number of rings"),
    ("  matrix size[1] ", "matrix size [1]"),
    ("Scaling Factor (mm/pixel) [2]", "scaling factor (mm/pixel) [2]"),
    ("!END OF INTERFILE", "end of interfile"),
    ("___", ""),
])
def test_keyword_standardisation(raw, std):
    assert run("key", raw).stdout.strip() == f"[{std}]"


def test_param_file_grammar(tmp_path):
    (tmp_path / "p.params").write_text(
        "\n  \n!OSEM_PARAMETERS :=\n"
        "; a comment line\n"
        "Input Projection File := ${FIR_TEST_DIR}/y.hs  \n"
        "scanner file:=  s.scan\n"
        "output volume header := ${FIR_MISSING}t.h33\n"
        "number of iterations := 3\n"
        "NUMBER_OF_SUBSETS := 16\n"
        "save interval :=\n"
        "cut radius in mm := 55.5\n"
        "convolution interval := 2\n"
        "convolution FHWM XYZ in mm := { 4 , 4.5,5}\n"
        "sensitivity map volume := sens.h33\n"
        "unknown key := 7\n"
        "!END OF OSEM PARAMETERS :=\n"
        "bias projection := after_stop.hs\n")
    env = dict(os.environ, FIR_TEST_DIR="/data")
    env.pop("FIR_MISSING", None)
    out = run("params", tmp_path / "p.params", env=env).stdout
    assert "'FIR_MISSING not found" in out
    vals = dict(l.split("=", 1) for l in out.strip().split("\n")
                if l.split("=", 1)[0].isalnum())
    assert vals["ok"] == "1"
    assert vals["in"] == "[/data/y.hs]"
    assert vals["scan"] == "[s.scan]"
    assert vals["hdr"] == "[t.h33]"
    assert vals["iter"] == "3" and vals["sub"] == "16"
    assert vals["save"] == "0"
    assert vals["cut"] == "55.5" and vals["conv"] == "2"
    assert vals["fwhm"] == "4,4.5,5,"
    assert vals["sens"] == "[sens.h33]" and vals["bias"] == "[]"


def test_param_file_without_start_key(tmp_path):
    (tmp_path / "p.params").write_text("number of iterations := 3\n")
    out = run("params", tmp_path / "p.params").stdout
    assert "ok=0" in out and "Required first keyword" in out


def test_missing_file_is_an_error(tmp_path):
    r = run("params", tmp_path / "nope.params", check=False)
    assert r.returncode != 0 and "Couldn't open file" in r.stderr


@pytest.mark.parametrize("name", ["c1", "c2", "c5"])
def test_python_scanner_tables_match_oracle(name, tmp_path):
    """fir_amd.lib.scanner_tables (which feeds the GPU) and the C++ host
    tables both equal the oracle's bit for bit (same libm sincos calls)."""
    from fir_amd import lib as firlib
    sc = synth.WORKLOADS[name]().scanner
    a, b = O.scanner_tables(sc)
    c, d = firlib.scanner_tables(sc)
    assert np.array_equal(a, c) and np.array_equal(b, d)
    sc.write(str(tmp_path / "s.scan"))
    out = run("scanner", tmp_path / "s.scan").stdout.split("\n")
    n, ns = map(int, out[0].split())
    xy = np.array([float(v) for line in out[1:1 + n] for v in line.split()])
    assert np.array_equal(xy, a)
