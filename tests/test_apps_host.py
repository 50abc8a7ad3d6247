"""Host-side checks of the drop-in executables and file formats (no GPU):
the .dat reader/writer against the oracle's C reader/writer and the
reference header layout, the C++ header compiling against the C ABI, and the
executables' argument and input-file errors, which happen before any device
call."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from gliomasolver_amd.datfile import read_dat, write_dat

BIN = os.path.join(ROOT, "gliomasolver_amd", "bin")


def test_dat_roundtrip_and_header(tmp_path, oracle):
    a = np.random.default_rng(0).uniform(size=(5, 6, 7))
    p = str(tmp_path / "a.dat")
    write_dat(p, a)
    raw = np.fromfile(p, np.int32, 6)
    assert list(raw) == [1234, 3, 7, 6, 5, 0]       # {magic, dim, nx, ny, nz}, type id double
    assert np.array_equal(read_dat(p), a)
    b = np.zeros_like(a)
    assert oracle.lib.gho_dat_read(p.encode(), 7, 6, 5, oracle.p(b)) == 0
    assert np.array_equal(a, b)


def test_dat_reads_float_and_2d(tmp_path):
    p = str(tmp_path / "f.dat")
    a = np.arange(24, dtype=np.float32).reshape(2, 3, 4)
    with open(p, "wb") as f:
        f.write(np.array([1234, 3, 4, 3, 2, 1], np.int32).tobytes())
        f.write(a.tobytes())
    assert np.array_equal(read_dat(p), a.astype(np.float64))
    roi = np.array([[1, 2, 3], [4, 5, 6]], np.float64)   # N x 3 points
    q = str(tmp_path / "roi.dat")
    write_dat(q, roi.T)                                   # (3, N): x fastest = point index
    assert list(np.fromfile(q, np.int32, 5)) == [1234, 2, 2, 3, 0]
    assert np.array_equal(read_dat(q).T, roi)


def test_executables_are_built():
    for name in ("brain", "likelihood", "rd_pipeline"):
        exe = os.path.join(BIN, name)
        assert os.access(exe, os.X_OK), exe


def test_hpp_compiles_standalone(tmp_path):
    src = tmp_path / "use.cpp"
    src.write_text('#include "gliomahip.hpp"\n'
                   'int main() { gliomahip::ReactionDiffusionOperator<3> r(1e-3, 1e-4, 0.02);\n'
                   '  gliomahip::UpdateTumor<3> u(0.1); return (int)(r.stencil_start[0] + u.stencil_end[0]); }\n')
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", "-I" + os.path.join(ROOT, "include"),
                        str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_brain_rejects_other_models(tmp_path):
    r = subprocess.run([os.path.join(BIN, "brain"), "-model", "PFFCH"], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 2 and "not provided" in r.stderr


def test_likelihood_requires_input_file(tmp_path):
    r = subprocess.run([os.path.join(BIN, "likelihood")], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 1 and "LikelihoodInput.txt" in r.stdout
