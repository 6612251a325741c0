"""Output writers against files written by the reference's own writers
(IO_VTK::WritePntsWithField_C, IO_AERECO::WriteVecs_C / ::WriteField_C;
recorded by tests/golden/make_golden.py), byte for byte: the Python writers
of bielcc_amd.io and the drop-in C++ header's IO_VTK / IO_AERECO."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from bielcc_amd import io as bio
from _util import GOLDEN

W = os.path.join(GOLDEN, "writers")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FILES = ["field_c.vtk", "vecs_real.gv", "vecs_image.gv", "field_real.gv", "field_image.gv", "field_abs.gdr"]


def _inputs():
    d = np.load(os.path.join(W, "inputs.npz"))
    return d["pts"], d["vals"]


def _same(a, b):
    with open(a, "rb") as fa, open(b, "rb") as fb:
        return fa.read() == fb.read()


def test_python_writers_match_reference_files(tmp_path):
    pts, vals = _inputs()
    bio.write_field_vtk_c(str(tmp_path / "field_c.vtk"), pts, vals)
    bio.write_aereco_vecs_c(str(tmp_path / "vecs_real.gv"), str(tmp_path / "vecs_image.gv"), vals)
    bio.write_aereco_field_c(str(tmp_path / "field_real.gv"), str(tmp_path / "field_image.gv"),
                             str(tmp_path / "field_abs.gdr"), vals[:4])
    for f in FILES:
        assert _same(tmp_path / f, os.path.join(W, f)), f


def test_field_writer_rejects_the_reference_overrun(tmp_path):
    _, vals = _inputs()
    with pytest.raises(ValueError):
        bio.write_aereco_field_c(str(tmp_path / "a"), str(tmp_path / "b"), str(tmp_path / "c"), vals)


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_cpp_writers_match_reference_files(tmp_path):
    pts, vals = _inputs()
    exe = tmp_path / "writers_check"
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), "-I",
                        os.path.join(ROOT, "bielcc_amd", "cpp"), os.path.join(ROOT, "tests", "cpp", "writers_check.cpp"),
                        "-o", str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    txt = tmp_path / "in.txt"
    with open(txt, "w") as f:
        f.write(f"{len(pts)}\n")
        for p in pts:
            f.write(" ".join("%.17g" % x for x in p) + "\n")
        for v in vals:
            f.write(" ".join("%.17g %.17g" % (c.real, c.imag) for c in v) + "\n")
    out = tmp_path / "out"
    out.mkdir()
    r = subprocess.run([str(exe), str(txt), str(out)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    for f in FILES:
        assert _same(out / f, os.path.join(W, f)), f
