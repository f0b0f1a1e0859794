"""The identities the GPU x pass and the Anderson Gram kernel rely on,
checked in FP64 NumPy (no GPU).

1. Three-line x pass (jaxmaterials_amd/csrc/ls_passes_reg.hpp, k_xpass_h): with
   xi~ = (alpha s_x, beta c_x, gamma c_x) along a line (s_x = sin(pi k / n),
   c_x = cos(pi k / n), k in [0, n) as in fourier.py:9-58), the divergence
   xi~ . sigma_hat equals e^{-i pi k/n} FFT[h] for the two-point stencil h, and
   IFFT[sym(xi~ (x) u)] equals the backward stencils of IFFT[e^{+i pi k/n} u].
2. Anderson Gram row by Parseval (ls_kernels.hpp, aa_y / k_aa_gram): per
   mode, Re(y . conj(y')) with y = sqrt(w_k / 2) |xi~| (u + (sqrt 2 - 1)(n . u) n)
   is w_k Re <sym(xi~ (x) u), sym(xi~ (x) u')>_w; summed over the z-half
   spectrum it is N times the Voigt-weighted real-space inner product of the
   compatible fields r = IFFT[sym(xi~ (x) u)] wherever the half spectrum is
   Hermitian-consistent.  On the self-conjugate planes k_z = 0 and k_z = nz/2
   the reference's xi~ (k in [0, n)) is odd under the in-plane flip, the C2R
   transform keeps only the Hermitian part there, and the spectral Gram row
   measures the pre-C2R spectrum instead: the test pins the identity with
   those planes empty and checks the per-mode algebra everywhere.
"""

import numpy as np

from oracle import ls_oracle as O

T = [[0, 3, 4], [3, 1, 5], [4, 5, 2]]  # tensor index -> Voigt component


def test_three_line_divergence_and_strain_identities():
    rng = np.random.default_rng(5)
    for n in (16, 15, 32):
        k = np.arange(n)
        sx, cx = np.sin(np.pi * k / n), np.cos(np.pi * k / n)
        if n % 2 == 0:
            cx[n // 2] = 0.0
        al, be, ga = 1.3, -0.7, 0.4
        xi = np.stack([al * sx, be * cx, ga * cx])
        sig = rng.normal(size=(6, n)) + 1j * rng.normal(size=(6, n))
        S = np.fft.fft(sig, axis=-1)
        d = np.stack([sum(xi[j] * S[T[i][j]] for j in range(3)) for i in range(3)])
        h = []
        for i in range(3):
            six = sig[T[i][0]]
            w = be * sig[T[i][1]] + ga * sig[T[i][2]]
            h.append(al / 2j * (np.roll(six, -1) - six) + 0.5 * (w + np.roll(w, -1)))
        ph = np.exp(-1j * np.pi * k / n)
        assert np.abs(ph * np.fft.fft(np.array(h), axis=-1) - d).max() <= 1e-12 * np.abs(d).max()

        u = rng.normal(size=(3, n)) + 1j * rng.normal(size=(3, n))
        g = np.fft.ifft(u / ph, axis=-1) * n  # unnormalised inverse of e^{+i pi k/n} u
        gm = np.roll(g, 1, axis=-1)
        D = lambda j: (g[j] - gm[j]) / 2j  # noqa: E731
        Av = lambda j: 0.5 * (g[j] + gm[j])  # noqa: E731
        coef = [al, be, ga]
        part = lambda i, j: coef[i] * (D(j) if i == 0 else Av(j))  # noqa: E731
        for i in range(3):
            for j in range(i, 3):
                r = np.fft.ifft(0.5 * (xi[i] * u[j] + xi[j] * u[i]), axis=-1) * n
                rr = 0.5 * (part(i, j) + part(j, i))
                assert np.abs(r - rr).max() <= 1e-12 * np.abs(r).max()


def test_anderson_gram_by_parseval():
    rng = np.random.default_rng(7)
    for shape, ext in [((8, 6, 10), (1.0, 0.7, 1.3)), ((7, 9, 5), (1.1, 0.9, 0.8))]:
        gs = O.GridSpec(*shape, *ext)
        N = np.prod(shape)
        xi = O.get_xi(gs, kz_half=True)
        q2 = np.sum(xi**2, axis=0)
        nz = shape[2]
        w = O.hermitian_weights(nz)[None, None, :] * np.ones(xi.shape[1:])

        def fields():
            uh = np.fft.rfftn(rng.normal(size=(3,) + shape), axes=(1, 2, 3))
            uh[:, :, :, 0] = 0.0  # the self-conjugate planes empty (see the module docstring)
            if nz % 2 == 0:
                uh[:, :, :, nz // 2] = 0.0
            rh = np.stack([0.5 * (xi[i] * uh[j] + xi[j] * uh[i])
                           for i, j in [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]])
            r = np.fft.irfftn(rh, s=shape, axes=(1, 2, 3))
            with np.errstate(invalid="ignore", divide="ignore"):
                inv = np.where(q2 > 0, 1.0 / np.sqrt(q2), 0.0)
            n = xi * inv
            pr = (np.sqrt(2.0) - 1.0) * np.sum(n * uh, axis=0)
            y = np.sqrt(0.5 * w) * np.sqrt(q2) * (uh + pr * n)
            return r, y, rh

        r1, y1, rh1 = fields()
        r2, y2, rh2 = fields()
        # per mode: Re(y . conj(y')) = w Re <r_hat, r_hat'>_w
        per_mode = np.real(np.einsum("aijk,aijk,a->ijk", rh1, np.conj(rh2), O.VOIGT_WEIGHTS)) * w
        assert np.abs(np.real(np.sum(y1 * np.conj(y2), axis=0)) - per_mode).max() <= 1e-10 * np.abs(per_mode).max()
        real_space = np.einsum("aijk,aijk,a->", r1, r2, O.VOIGT_WEIGHTS)
        spectral = np.sum(np.real(y1 * np.conj(y2))) / N
        assert abs(real_space - spectral) <= 1e-10 * abs(real_space)
        self_real = np.einsum("aijk,aijk,a->", r1, r1, O.VOIGT_WEIGHTS)
        assert abs(self_real - np.sum(np.abs(y1) ** 2) / N) <= 1e-10 * self_real
