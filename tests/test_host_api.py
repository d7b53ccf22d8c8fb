"""Source compatibility of the kept C++ interfaces: tests/host_api/caller.cc,
a program that uses only the reference's public ScannerData/VolData API
(ScannerData::getCrystalCoordinates and printBinInfo, ScannerData.h:29-43;
VolData::computeLineIntegral and projectLineIntegral, VolData.h:108-112),
compiled against fir_amd/host/include and linked with libfir_host.so, must
print exactly what the same source prints when compiled against the
reference (tests/golden/host_api_golden.json). Host only: no GPU call."""
import json
import os
import shutil
import subprocess

import pytest

from conftest import GOLDEN, ROOT
from fir_amd import interfile as io
from fir_amd import synth

HOST = os.path.join(ROOT, "fir_amd", "host", "include")


def test_reference_style_caller_compiles_and_matches(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("no host C++ compiler")
    lib = os.path.join(ROOT, "fir_amd", "libfir_host.so")
    if not os.path.exists(lib):
        pytest.fail("host layer not built (python -m fir_amd.build)")
    g = json.load(open(os.path.join(GOLDEN, "host_api_golden.json")))
    exe = str(tmp_path / "caller")
    subprocess.run(
        ["g++", "-O2", "-std=c++17", f"-I{HOST}",
         f"-I{os.path.join(ROOT, 'include')}", "-o", exe,
         os.path.join(ROOT, "tests", "host_api", "caller.cc"),
         f"-L{os.path.join(ROOT, 'fir_amd')}", "-lfir_host", "-lfir_hip",
         f"-Wl,-rpath,{os.path.join(ROOT, 'fir_amd')}"],
        check=True, capture_output=True, text=True)
    io.ScannerSpec(**g["scanner"]).write(str(tmp_path / "s.scan"))
    w = synth.c1()
    w.vol.write(str(tmp_path / "v.h33"), synth.phantom(w)[0])
    r = subprocess.run([exe, "s.scan", "v.h33", *g["args"]], cwd=tmp_path,
                       capture_output=True, text=True,
                       env=dict(os.environ, FIR_PINNED="0"))
    assert r.returncode == 0, r.stderr
    assert r.stdout == g["stdout"]
