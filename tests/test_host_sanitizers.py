"""Host C++ of the library under AddressSanitizer and UndefinedBehaviorSanitizer
(g++, CPU only): the fibre movement (jaxmaterials_amd/csrc/fibres.cpp, driven
by tests/host_dft/fibres_sanitize.cpp), the register DFTs
(tests/host_dft/dft_check.cpp) and the damage split
(tests/host_dft/damage_check.cpp).  Any report aborts the program
(-fno-sanitize-recover=all), which fails the test."""

import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]


def _gxx():
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    return gxx


def _build(tmp_path, name, sources):
    exe = str(tmp_path / name)
    subprocess.run([_gxx(), "-std=c++17"] + SAN + ["-I", os.path.join(HERE, "host_dft"), "-o", exe] + sources,
                   check=True)
    return exe


def _run(exe, stdin=None):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    p = subprocess.run([exe], input=stdin, capture_output=True, text=True, env=env, timeout=300)
    assert p.returncode == 0, p.stderr[-4000:]
    assert "runtime error" not in p.stderr and "AddressSanitizer" not in p.stderr, p.stderr[-4000:]
    return p.stdout


def test_fibre_movement_under_sanitizers(tmp_path):
    exe = _build(tmp_path, "fibres_san", [os.path.join(HERE, "host_dft", "fibres_sanitize.cpp"),
                                          os.path.join(ROOT, "jaxmaterials_amd", "csrc", "fibres.cpp")])
    out = _run(exe).split()
    assert len(out) == 4 and all(float(v) >= 0.0 for v in out)


def test_register_dfts_under_sanitizers(tmp_path):
    exe = _build(tmp_path, "dft_san", [os.path.join(HERE, "host_dft", "dft_check.cpp")])
    rows = _run(exe).strip().splitlines()
    assert len(rows) == 10


def test_damage_split_under_sanitizers(tmp_path):
    exe = _build(tmp_path, "damage_san", [os.path.join(HERE, "host_dft", "damage_check.cpp")])
    lines = "\n".join(["2.593 1.111 0.01 -0.02 0.003 0.001 0 0", "19.44 29.17 -0.01 -0.01 -0.01 0 0 0",
                       "1 1 0 0 0 0.5 0.5 0.5", "1 1 0.1 0.1 0.1 0 0 0"]) + "\n"
    out = _run(exe, stdin=lines).strip().splitlines()
    assert len(out) == 4
