"""Register DFTs (jaxmaterials_amd/csrc/fft_reg.hpp) compiled as host code and
checked against the DFT definition for every length the GPU kernels use
(2, 3, 4, 5, 6, 8, 12, 16, 24, 32; forward and inverse).  The lengths 6-32
are compile-time Cooley-Tukey splits with constant twiddles; a wrong twiddle
table entry or split index shows up as an O(1) error here."""

import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def dft_rows(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("dft") / "dft_check")
    subprocess.run([gxx, "-std=c++17", "-O1", "-I", os.path.join(HERE, "host_dft"), "-o", exe,
                    os.path.join(HERE, "host_dft", "dft_check.cpp")], check=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    return {int(r): (float(f), float(i)) for r, f, i in (line.split() for line in out.strip().splitlines())}


def test_every_register_dft_length(dft_rows):
    assert sorted(dft_rows) == [2, 3, 4, 5, 6, 8, 12, 16, 24, 32]
    for r, (fwd, inv) in dft_rows.items():
        # FP32 butterflies: a few ulp relative to the largest output
        assert fwd < 5e-7 and inv < 5e-7, (r, fwd, inv)
