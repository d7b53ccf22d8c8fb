"""The drop-in executables (fir_amd/bin/FIR_ForwardProj, FIR_BackProj,
FIR_OSEM) on the 4-ring x 128-crystal workload: same parameter files and
outputs as the reference executables, results checked against the
reference's own outputs (tests/golden)."""
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, rel_l2, rel_max
from fir_amd import interfile as io
from fir_amd import synth

pytestmark = pytest.mark.gpu
BIN = os.path.join(ROOT, "fir_amd", "bin")


def exe(name, *args, cwd, check=True, env=None):
    r = subprocess.run([os.path.join(BIN, name), *map(str, args)], cwd=cwd,
                       capture_output=True, text=True, timeout=120,
                       env=None if env is None else dict(os.environ, **env))
    if check and r.returncode != 0:
        raise AssertionError(f"{name} failed:\n{r.stdout[-2000:]}\n{r.stderr}")
    return r


@pytest.fixture(scope="module")
def work(tmp_path_factory, c1_golden):
    d = tmp_path_factory.mktemp("cli")
    w = synth.c1()
    w.scanner.write(str(d / "scanner.scan"))
    w.proj.write(str(d / "tmpl.hs"))
    w.vol.write(str(d / "tmpl.h33"))
    w.vol.write(str(d / "act.h33"), c1_golden["act"])
    w.vol.write(str(d / "hu.h33"), c1_golden["hu"])
    io.write_proj(str(d / "y.hs"), w.proj, c1_golden["y"])
    io.write_proj(str(d / "bias.hs"), w.proj,
                  np.full_like(c1_golden["y"], float(c1_golden["bias_value"])))
    return d


@pytest.fixture(scope="module")
def headers():
    return json.load(open(os.path.join(GOLDEN, "ref_headers.json")))


def osem_params(path, out, n_iter, n_sub, **kw):
    items = [("input projection file", "y.hs"),
             ("scanner file", "scanner.scan"),
             ("output volume header", "tmpl.h33"),
             ("output volume file name", out),
             ("number of iterations", str(n_iter)),
             ("number of subsets", str(n_sub))]
    extra = {k.replace("_", " "): v for k, v in kw.items()}
    items = [(k, extra.pop(k, v)) for k, v in items] + list(extra.items())
    io.write_params(str(path), "OSEM PARAMETERS", items)


def test_forward_proj_executable(work, c1_golden, headers):
    io.write_params(str(work / "fwd.params"), "FORWARD PROJECTION PARAMETERS",
                    [("input volume file", "act.h33"),
                     ("scanner file", "scanner.scan"),
                     ("output projection header", "tmpl.hs"),
                     ("output projection file name", "y0")])
    exe("FIR_ForwardProj", "fwd.params", cwd=work)
    assert (work / "y0.hs").read_text() == headers["y0.hs"]
    got = np.fromfile(work / "y0.s", np.float32)
    assert np.array_equal(got, c1_golden["y0"])


def test_back_proj_executable(work, c1_golden, headers):
    exe("FIR_BackProj", "y.hs", "scanner.scan", "tmpl.h33", "bp", cwd=work)
    assert (work / "bp.h33").read_text() == headers["bp.h33"]
    got = np.fromfile(work / "bp.i33", np.float32)
    assert rel_max(got, c1_golden["bp"]) < 1e-5


def test_back_proj_requires_four_arguments(work):
    r = exe("FIR_BackProj", "y.hs", "scanner.scan", "tmpl.h33", cwd=work,
            check=False)
    assert r.returncode != 0 and "four input arguments" in r.stderr


def test_osem_plain(work, c1_golden, headers):
    osem_params(work / "a.params", "xa", 2, 4,
                sensitivity_map_volume="sens.h33", save_interval="3")
    exe("FIR_OSEM", "a.params", cwd=work)
    assert (work / "xa.h33").read_text() == headers["xa.h33"]
    assert (work / "sens.h33").read_text() == headers["sens.h33"]
    xa = np.fromfile(work / "xa.i33", np.float32)
    assert rel_l2(xa, c1_golden["xa"]) < 1e-4
    sens = np.fromfile(work / "sens.i33", np.float32)
    assert rel_max(sens, c1_golden["sens"]) < 1e-5
    # intermediate saves every 3 sub-iterations, not the last one
    assert (work / "xa_subiter_3.h33").exists()
    assert (work / "xa_subiter_6.h33").exists()
    assert not (work / "xa_subiter_8.h33").exists()
    # reuse of the saved sensitivity map (flag 0) gives the same image
    osem_params(work / "a2.params", "xa2", 2, 4,
                sensitivity_map_volume="sens.h33")
    r = exe("FIR_OSEM", "a2.params", "0", cwd=work)
    assert "Reading sensitivity map from file" in r.stdout
    xa2 = np.fromfile(work / "xa2.i33", np.float32)
    assert rel_l2(xa2, xa) < 1e-6


def test_osem_attenuation_gaussian_cut(work, c1_golden):
    osem_params(work / "b.params", "xb", 2, 4,
                attenuation_volume_in_HU="hu.h33",
                attenuation_correction_factors="acf.hs",
                cut_radius_in_mm="55", convolution_interval="1",
                convolution_FHWM_XYZ_in_mm="{4,4,4}")
    exe("FIR_OSEM", "b.params", cwd=work)
    xb = np.fromfile(work / "xb.i33", np.float32)
    assert rel_l2(xb, c1_golden["xb"]) < 1e-4
    assert (work / "acf.hs").exists()
    # ACF reused from file when the recompute flag is 0
    osem_params(work / "b2.params", "xb2", 2, 4,
                attenuation_volume_in_HU="hu.h33",
                attenuation_correction_factors="acf.hs",
                cut_radius_in_mm="55", convolution_interval="1",
                convolution_FHWM_XYZ_in_mm="{4,4,4}")
    r = exe("FIR_OSEM", "b2.params", "1", "0", cwd=work)
    assert "Reading attenuation correction factors" in r.stdout
    xb2 = np.fromfile(work / "xb2.i33", np.float32)
    assert rel_l2(xb2, xb) < 1e-6


def test_osem_bias(work, c1_golden):
    osem_params(work / "c.params", "xc", 1, 4, bias_projection="bias.hs")
    exe("FIR_OSEM", "c.params", cwd=work)
    xc = np.fromfile(work / "xc.i33", np.float32)
    assert rel_l2(xc, c1_golden["xc"]) < 1e-4


@pytest.mark.parametrize("fmt", ["snorm16", "fp32"])
def test_osem_path_cache(work, c1_golden, fmt):
    """FIR_PATH_CACHE selects the path-cache projector: runs A and B of the
    reference within 1e-4 relative L2."""
    osem_params(work / f"pa_{fmt}.params", f"pa_{fmt}", 2, 4)
    r = exe("FIR_OSEM", f"pa_{fmt}.params", cwd=work,
            env={"FIR_PATH_CACHE": fmt})
    assert "Building the path cache" in r.stdout
    xa = np.fromfile(work / f"pa_{fmt}.i33", np.float32)
    assert rel_l2(xa, c1_golden["xa"]) < 1e-4
    osem_params(work / f"pb_{fmt}.params", f"pb_{fmt}", 2, 4,
                attenuation_volume_in_HU="hu.h33",
                cut_radius_in_mm="55", convolution_interval="1",
                convolution_FHWM_XYZ_in_mm="{4,4,4}")
    exe("FIR_OSEM", f"pb_{fmt}.params", cwd=work, env={"FIR_PATH_CACHE": fmt})
    xb = np.fromfile(work / f"pb_{fmt}.i33", np.float32)
    assert rel_l2(xb, c1_golden["xb"]) < 1e-4


def test_osem_partial_path_cache(work, c1_golden):
    """FIR_PATH_CACHE_SUBSETS=2: subsets 1-2 of 4 from a partial cache, 3-4
    walked on the fly; run B of the reference within 1e-4 relative L2. A
    budget too small for one subset fails with the cache's message."""
    osem_params(work / "pp.params", "pp", 2, 4,
                attenuation_volume_in_HU="hu.h33",
                cut_radius_in_mm="55", convolution_interval="1",
                convolution_FHWM_XYZ_in_mm="{4,4,4}")
    r = exe("FIR_OSEM", "pp.params", cwd=work,
            env={"FIR_PATH_CACHE": "snorm16", "FIR_PATH_CACHE_SUBSETS": "2"})
    assert "holds subsets 1 to 2 of 4" in r.stdout
    xb = np.fromfile(work / "pp.i33", np.float32)
    assert rel_l2(xb, c1_golden["xb"]) < 1e-4
    r = exe("FIR_OSEM", "pp.params", cwd=work, check=False,
            env={"FIR_PATH_CACHE": "snorm16", "FIR_PATH_CACHE_GB": "1e-9"})
    assert r.returncode != 0 and "subset 0 alone needs" in r.stderr


@pytest.mark.parametrize("cache", ["off", "snorm16"])
@pytest.mark.parametrize("run", ["xd", "xe"])
def test_osem_reso_reco(work, run, cache):
    """FIR_RESO_RECO=1 selects reconAlgos::OSEM_ResoReco (a constant false
    in the reference's OSEM.cc:340-341): runs D and E of the compiled
    reference's OSEM_ResoReco within 1e-4 relative L2, on the plain and the
    cached projector."""
    g = np.load(os.path.join(GOLDEN, "reso_golden.npz"))
    nit, ns, fw, conv, cut = g[run + "_params"]
    name = f"r_{run}_{cache}"
    osem_params(work / f"{name}.params", name, int(nit), int(ns),
                convolution_interval=str(int(conv)),
                convolution_FHWM_XYZ_in_mm=f"{{{fw},{fw},{fw}}}",
                cut_radius_in_mm=str(float(cut)))
    r = exe("FIR_OSEM", f"{name}.params", cwd=work,
            env={"FIR_RESO_RECO": "1", "FIR_PATH_CACHE": cache})
    assert "OSEM_ResoReco" in r.stdout
    x = np.fromfile(work / f"{name}.i33", np.float32)
    assert rel_l2(x, g[run]) < 1e-4


def test_host_containers_use_pinned_memory(work, c1_golden):
    """Sinograms and volumes of the host layer live in page-locked memory
    when a GPU is present (FIR_PINNED=0 keeps ordinary memory); transfers
    and results are the same either way."""
    sel = os.path.join(BIN, "fir_host_selftest")
    outs = {}
    for flag in ("1", "0"):
        r = subprocess.run([sel, "upload", "y.hs", "3"], cwd=work,
                           capture_output=True, text=True, timeout=120,
                           env=dict(os.environ, FIR_PINNED=flag))
        assert r.returncode == 0, r.stderr
        outs[flag] = json.loads(r.stdout.strip().splitlines()[-1])
    assert outs["1"]["pinned"] == 1 and outs["0"]["pinned"] == 0
    assert outs["1"]["bytes"] == 4 * len(c1_golden["y"])
    osem_params(work / "pp.params", "pp", 1, 4)
    exe("FIR_OSEM", "pp.params", cwd=work, env={"FIR_PINNED": "0"})
    exe("FIR_OSEM", "pp.params", cwd=work)
    a = np.fromfile(work / "pp.i33", np.float32)
    assert np.isfinite(a).all() and a.max() > 0


def test_osem_path_cache_rejects_unknown_format(work):
    osem_params(work / "pz.params", "pz", 1, 4)
    r = exe("FIR_OSEM", "pz.params", cwd=work, check=False,
            env={"FIR_PATH_CACHE": "int8"})
    assert r.returncode != 0
    assert "FIR_PATH_CACHE" in r.stderr


def test_osem_rejects_non_divisor_subsets(work):
    osem_params(work / "bad.params", "xbad", 1, 3)
    r = exe("FIR_OSEM", "bad.params", cwd=work, check=False)
    assert r.returncode != 0
    assert "Number of subsets must be a divisor of 64" in r.stderr


def test_forward_proj_rejects_bad_frame(work):
    """A frame index outside the volume is an error (ForwardProj.cc:170-178)."""
    io.write_params(str(work / "fbad.params"), "FORWARD PROJECTION PARAMETERS",
                    [("input volume file", "act.h33"),
                     ("scanner file", "scanner.scan"),
                     ("output projection header", "tmpl.hs"),
                     ("output projection file name", "yb"),
                     ("frames to project", "{3}")])
    r = exe("FIR_ForwardProj", "fbad.params", cwd=work, check=False)
    assert r.returncode != 0
    assert "Frame indices must be between 0 and 0" in r.stderr


def test_osem_rejects_scanner_projection_mismatch(work):
    """A sinogram of another crystal count than the scanner's is an error
    (ScannerData.cc:31-44)."""
    import dataclasses
    spec = dataclasses.replace(synth.c1().proj, n_crystals_per_ring=64,
                               n_tang_coords=33)
    n = sum(4 - abs(s) for s in range(-3, 4)) * 32 * 33
    io.write_proj(str(work / "y64.hs"), spec, np.zeros(n, np.float32))
    osem_params(work / "mm.params", "xmm", 1, 4,
                input_projection_file="y64.hs")
    r = exe("FIR_OSEM", "mm.params", cwd=work, check=False)
    assert r.returncode != 0
    assert "same number of crystals per ring (128)" in r.stderr


def test_osem_rejects_mismatched_sensitivity_file(work):
    """A sensitivity map file (read, recompute flag 0) whose header differs
    from the output volume's is an error (OSEM.cc:238-242)."""
    import dataclasses
    vol = dataclasses.replace(synth.c1().vol, size=(32, 32, 8), n_frames=4)
    vol.write(str(work / "sens32.h33"),
              np.ones(4 * 32 * 32 * 8, np.float32))
    osem_params(work / "sm.params", "xsm", 1, 4,
                sensitivity_map_volume="sens32.h33")
    r = exe("FIR_OSEM", "sm.params", "0", cwd=work, check=False)
    assert r.returncode != 0
    assert "doesn't fit with output volume" in r.stderr


def test_missing_parameter_file(work):
    r = exe("FIR_OSEM", cwd=work, check=False)
    assert r.returncode != 0 and "Parameter file missing" in r.stderr
    r = exe("FIR_ForwardProj", "nope.params", cwd=work, check=False)
    assert r.returncode != 0 and "Couldn't open file" in r.stderr
