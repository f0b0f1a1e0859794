"""The damage split of the phase-field law (jaxmaterials_amd/csrc/damage.hpp:
closed-form eigenvalues and spectral projectors) compiled as host code and
checked against the oracle (numpy.linalg.eigh) on random strains and on the
degenerate cases: repeated eigenvalues, eigenvalues at zero, pure shear,
zero strain, all-positive and all-negative spectra."""

import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import phasefield_oracle as PF

HERE = os.path.dirname(os.path.abspath(__file__))


def _cases():
    rng = np.random.default_rng(11)
    rows = []
    for _ in range(400):
        rows.append((rng.uniform(0.1, 3), rng.uniform(0.1, 3), *rng.normal(0, 1, 6)))
    Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    for diag in ([1.0, 1.0, -2.0], [-1.0, -1.0, 2.0], [0.0, 0.0, 1.0], [0.0, -1.0, 1.0], [2.0, 2.0, 2.0],
                 [-3.0, -3.0, -3.0], [0.0, 0.0, 0.0], [1e-9, -1e-9, 0.5], [1.0, 1.0 + 1e-12, -1.0]):
        m = Q @ np.diag(diag) @ Q.T
        rows.append((1.3, 0.7, m[0, 0], m[1, 1], m[2, 2], m[1, 2], m[0, 2], m[0, 1]))
    rows.append((1.0, 1.0, 0, 0, 0, 0.4, 0, 0))  # pure shear in one slot
    rows.append((1.0, 1.0, 0.5, 0.2, 0.1, 0, 0, 0))  # diagonal, all positive
    return np.array(rows)


@pytest.fixture(scope="module")
def host_rows():
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    import tempfile

    exe = os.path.join(tempfile.mkdtemp(), "damage_check")
    subprocess.run([gxx, "-std=c++17", "-O2", "-I", os.path.join(HERE, "host_dft"), "-o", exe,
                    os.path.join(HERE, "host_dft", "damage_check.cpp")], check=True)
    cases = _cases()
    inp = "\n".join(" ".join(f"{x:.17g}" for x in row) for row in cases) + "\n"
    out = subprocess.run([exe], input=inp, check=True, capture_output=True, text=True).stdout
    return cases, np.array([[float(t) for t in line.split()] for line in out.strip().splitlines()])


def test_tension_stress_and_energy_match_the_oracle(host_rows):
    cases, got = host_rows
    lam, mu, eps = cases[:, 0], cases[:, 1], cases[:, 2:].T
    # sigma+ = sigma(d = 0, k = 1) - Hooke, since g = 2 there
    splus = PF.sigma_damaged(lam, mu, eps, np.zeros_like(lam), 1.0) - PF.sigma_damaged(lam, mu, eps)
    scale = np.abs(splus).max(axis=0) + np.abs(eps).max(axis=0) * (lam + 2 * mu) + 1e-300
    err = np.abs(got[:, :6].T - splus).max(axis=0) / scale
    # generic spectra to FP64 rounding; the rotated matrices with a repeated
    # eigenvalue (the last rows) to the closed form's ~1e-8 resolution of the pair
    assert err[:400].max() < 1e-12, err[:400].max()
    assert err.max() < 1e-8, err.max()
    # exact on axis-aligned states
    assert err[-1] == 0.0
    psi = PF.strain_energy_plus(lam, mu, eps)
    assert np.allclose(got[:, 6], psi, rtol=1e-12, atol=1e-12 * (np.abs(eps).max() ** 2))
