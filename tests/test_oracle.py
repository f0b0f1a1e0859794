"""CPU pins of the FP64 oracle (oracle/ls_oracle.py).

Each test checks an oracle function against something that does not reuse its
formula: the reference's own tests and known answers, closed forms, an
independent tensor-index construction, brute force on tiny grids, or a
textbook equivalence.
"""

import numpy as np
import pytest

from oracle import ls_oracle as O

PY_GRIDS = {"even": (64, 48, 32), "odd": (57, 43, 37)}


def py_grid(name):
    return O.GridSpec(*PY_GRIDS[name], 1.2, 0.8, 0.9)


def material(shape, rng, lo=0.8, hi=1.2):
    mu = rng.uniform(lo, hi, size=shape)
    lm = rng.uniform(lo, hi, size=shape)
    return lm, mu


# --------------------------------------------------------------------------- D-
def test_backward_derivative_matches_stencil_loop():
    """D- against the explicit per-voxel index formula of derivatives.cu:9-93."""
    gs = O.GridSpec(5, 4, 3, 1.3, 0.7, 0.9)
    rng = np.random.default_rng(1)
    u = rng.normal(size=(gs.nx, gs.ny, gs.nz))
    nx, ny, nz = gs.shape
    for d, (n, L) in enumerate([(nx, gs.Lx), (ny, gs.Ly), (nz, gs.Lz)]):
        ref = np.zeros_like(u)
        for a in range(nx):
            for b in range(ny):
                for c in range(nz):
                    am, bm, cm = (a - 1) % nx, (b - 1) % ny, (c - 1) % nz
                    if d == 0:
                        plus = u[a, b, c] + u[a, bm, c] + u[a, b, cm] + u[a, bm, cm]
                        minus = u[am, b, c] + u[am, bm, c] + u[am, b, cm] + u[am, bm, cm]
                    elif d == 1:
                        plus = u[a, b, c] + u[am, b, c] + u[a, b, cm] + u[am, b, cm]
                        minus = u[a, bm, c] + u[am, bm, c] + u[a, bm, cm] + u[am, bm, cm]
                    else:
                        plus = u[a, b, c] + u[am, b, c] + u[a, bm, c] + u[am, bm, c]
                        minus = u[a, b, cm] + u[am, b, cm] + u[a, bm, cm] + u[am, bm, cm]
                    ref[a, b, c] = 0.25 * n / L * (plus - minus)
        assert np.allclose(O.backward_derivative(u, gs, d), ref, rtol=1e-13, atol=1e-13)


# --------------------------------------------------------------------------- xi
@pytest.mark.parametrize("grid", ["even", "odd"])
def test_xizero_exact_zero_set_matches_reference_threshold(grid):
    gs = py_grid(grid)
    a = O.get_xizero(gs)
    b = O.get_xizero_thresholded(gs)
    assert np.abs(a - b).max() < 1e-12
    xi = O.get_xi(gs)
    nrm = np.sqrt((xi**2).sum(0))
    zero = nrm == 0
    # zero set: k = 0, or at least two axes at their Nyquist index
    K = np.meshgrid(*[np.arange(n) for n in gs.shape], indexing="ij")
    nyq = sum(((K[d] * 2 == n) & (n % 2 == 0)).astype(int) for d, n in enumerate(gs.shape))
    expect = ((K[0] == 0) & (K[1] == 0) & (K[2] == 0)) | (nyq >= 2)
    assert np.array_equal(zero, expect)


def test_xi_closed_form():
    gs = O.GridSpec(6, 5, 4, 1.5, 2.0, 0.5)
    xi = O.get_xi(gs)
    k = (2, 3, 1)
    h = [np.pi * k[d] / gs.shape[d] for d in range(3)]
    L = [gs.Lx, gs.Ly, gs.Lz]
    for d in range(3):
        v = 2 * gs.shape[d] / L[d] * np.sin(h[d])
        for e in range(3):
            if e != d:
                v *= np.cos(h[e])
        assert abs(xi[d][k] - v) < 1e-13


# --------------------------------------------------------------------------- Gamma0
def voigt_to_tensor(v):
    t = np.empty((3, 3) + v.shape[1:], dtype=v.dtype)
    t[0, 0], t[1, 1], t[2, 2] = v[0], v[1], v[2]
    t[0, 1] = t[1, 0] = v[3]
    t[0, 2] = t[2, 0] = v[4]
    t[1, 2] = t[2, 1] = v[5]
    return t


def tensor_to_voigt(t):
    return np.stack([t[0, 0], t[1, 1], t[2, 2], t[0, 1], t[0, 2], t[1, 2]])


def gamma0_tensor(xi0, lam0, mu0):
    """Gamma0_{klij} of doc/linear_elasticity.tex:137-152 built index by index."""
    d = np.eye(3)
    A = 0.25 * (np.einsum("li,j,k->klij", d, xi0, xi0) + np.einsum("ki,j,l->klij", d, xi0, xi0)
                + np.einsum("lj,i,k->klij", d, xi0, xi0) + np.einsum("kj,i,l->klij", d, xi0, xi0))
    B = np.einsum("i,j,k,l->klij", xi0, xi0, xi0, xi0)
    return (A - (lam0 + mu0) / (lam0 + 2 * mu0) * B) / mu0


def test_fourier_solve_matches_tensor_construction():
    rng = np.random.default_rng(5)
    lam0, mu0 = 1.7, 0.9
    for _ in range(20):
        x = rng.normal(size=3)
        x /= np.linalg.norm(x)
        tau = rng.normal(size=6) + 1j * rng.normal(size=6)
        G = gamma0_tensor(x, lam0, mu0)
        ref = -tensor_to_voigt(np.einsum("klij,ij->kl", G, voigt_to_tensor(tau)))
        got = O.fourier_solve(tau, lam0, mu0, x)
        assert np.allclose(got, ref, rtol=1e-12, atol=1e-12)


def test_gamma0_projector_and_symmetry():
    """Gamma0 C0 Gamma0 = Gamma0 per mode; Gamma0 symmetric in the tensor (Voigt-weighted) product."""
    rng = np.random.default_rng(6)
    lam0, mu0 = 0.8, 1.3
    W = O.VOIGT_WEIGHTS
    for _ in range(20):
        x = rng.normal(size=3)
        x /= np.linalg.norm(x)
        G = np.stack([-O.fourier_solve(e, lam0, mu0, x) for e in np.eye(6)], axis=1)  # Gamma0 as 6x6
        C0 = np.zeros((6, 6))
        C0[:3, :3] = lam0
        C0 += 2 * mu0 * np.eye(6)
        assert np.allclose(G @ C0 @ G, G, atol=1e-12)
        assert np.allclose(np.diag(W) @ G, (np.diag(W) @ G).T, atol=1e-12)


@pytest.mark.parametrize("grid", ["even", "odd"])
def test_div_sigma_invariant(grid):
    """TestDivSigma (cuda/src/tests/test_fourier.hh:267-316) in FP64: sigma = C0 eps + tau
    with eps = -Gamma0 tau is divergence free."""
    gs = O.GridSpec(*{"even": (48, 64, 32), "odd": (47, 61, 37)}[grid], 1.1, 0.9, 0.7)
    rng = np.random.default_rng(7812481)
    tau = rng.normal(size=(6,) + gs.shape)
    lam0, mu0 = 0.8, 1.3
    eps = np.real(np.fft.ifftn(O.fourier_solve(np.fft.fftn(tau, axes=(1, 2, 3)), lam0, mu0, O.get_xizero(gs)),
                               axes=(1, 2, 3)))
    sig = tau + 2 * mu0 * eps
    sig[:3] += lam0 * (eps[0] + eps[1] + eps[2])
    div = O.backward_divergence(sig, gs)
    rel = np.sqrt((div**2).sum()) / np.sqrt((O.VOIGT_WEIGHTS[:, None, None, None] * sig**2).sum())
    assert rel < 1e-12


# --------------------------------------------------------------------------- residual
@pytest.mark.parametrize("grid", ["even", "odd"])
def test_relative_convergence(grid):
    """test_relative_convergence (tests/test_lippmann_schwinger.py:50-61)."""
    gs = py_grid(grid)
    xi = O.get_xi(gs)
    rng = np.random.default_rng(seed=8741823)
    sigma = rng.normal(size=(6,) + gs.shape)
    sigma_hat = np.fft.fftn(sigma, axes=[-3, -2, -1])
    r_real = O.relative_divergence(sigma, gs)
    r_four = O.relative_divergence_fourier(sigma_hat, xi, gs)
    assert abs((r_real - r_four) / r_real) < 1.0e-5
    # the half-spectrum form with Hermitian weights gives the same number
    r_half, _ = O.relative_divergence_rfft(np.fft.rfftn(sigma, axes=(1, 2, 3)), O.get_xi(gs, kz_half=True), gs.nz)
    assert abs(r_half - r_four) / r_four < 1e-12


# --------------------------------------------------------------------------- whole solve
@pytest.mark.parametrize("grid", ["even", "odd"])
@pytest.mark.parametrize("depth", [0, 2, 4])
def test_convergence_reference_bounds(grid, depth):
    """test_convergence (tests/test_lippmann_schwinger.py:64-99), float64 branch."""
    gs = py_grid(grid)
    rng = np.random.default_rng(seed=784173)
    shape = gs.shape
    mu = np.ones(shape) + rng.uniform(size=shape, low=-0.2, high=+0.2)
    lmbda = np.ones(shape) + rng.uniform(size=shape, low=-0.2, high=+0.2)
    eb = np.array([2.1, 0.9, 0.8, 0.4, 0.9, 0.5])
    # reference quirk: initialise_material returns (mu, lambda) bound as (lmbda, mu)
    _, sigma, it, _ = O.lippmann_schwinger_rfft(mu, lmbda, eb, gs, rtol=1e-20, atol=1e-12, depth=depth, maxiter=32)
    assert it < {0: 16, 2: 14, 4: 13}[depth]
    assert O.relative_divergence(sigma, gs) < 1e-12


def test_full_and_half_spectrum_restatements_agree():
    gs = O.GridSpec(12, 10, 9, 1.0, 1.1, 0.9)
    rng = np.random.default_rng(3)
    lm, mu = material(gs.shape, rng, 0.5, 3.0)
    eb = rng.normal(size=6)
    for depth in (0, 3):
        e1, s1, i1, _ = O.lippmann_schwinger_jax(lm, mu, eb, gs, rtol=1e-10, depth=depth, maxiter=200)
        e2, s2, i2, _ = O.lippmann_schwinger_rfft(lm, mu, eb, gs, rtol=1e-10, depth=depth, maxiter=200)
        assert i1 == i2
        assert np.abs(e1 - e2).max() < 1e-12


def test_homogeneous_zero_iterations():
    """TestHomogeneousMaterial (cuda/src/tests/test_lippmann_schwinger.hh:104-128)."""
    gs = O.GridSpec(8, 6, 5, 1.1, 0.9, 0.7)
    lam = np.full(gs.shape, 1.2)
    eb = np.array([1.0, 0.4, 0.3, 0.1, -0.4, 0.7])
    eps, sig, it = O.lippmann_schwinger_capi(lam, lam, eb, gs, 1e-20, 1e-4, 100)
    assert it == 0
    assert np.allclose(eps, eb[:, None, None, None])


def dense_solution(lm, mu, eb, gs):
    """Brute force: assemble D- (derivatives.py), D+ = -(D-)^T, strain = eps_bar + sym(D+ u),
    solve D- . (C strain) = 0 densely (least squares; the strain is unique)."""
    N = gs.number_of_voxels
    I = np.eye(N).reshape((N,) + gs.shape)
    Dm = [np.stack([O.backward_derivative(I[i], gs, d).ravel() for i in range(N)], axis=1) for d in range(3)]
    Dp = [-D.T for D in Dm]
    Z = np.zeros((N, N))
    # strain (6N) from u (3N)
    B = np.block([
        [Dp[0], Z, Z],
        [Z, Dp[1], Z],
        [Z, Z, Dp[2]],
        [0.5 * Dp[1], 0.5 * Dp[0], Z],
        [0.5 * Dp[2], Z, 0.5 * Dp[0]],
        [Z, 0.5 * Dp[2], 0.5 * Dp[1]],
    ])
    l, m = np.diag(lm.ravel()), np.diag(mu.ravel())
    C = np.block([[l + 2 * m, l, l, Z, Z, Z], [l, l + 2 * m, l, Z, Z, Z], [l, l, l + 2 * m, Z, Z, Z],
                  [Z, Z, Z, 2 * m, Z, Z], [Z, Z, Z, Z, 2 * m, Z], [Z, Z, Z, Z, Z, 2 * m]])
    Dv = np.block([[Dm[0], Z, Z, Dm[1], Dm[2], Z], [Z, Dm[1], Z, Dm[0], Z, Dm[2]], [Z, Z, Dm[2], Z, Dm[0], Dm[1]]])
    K = Dv @ C @ B
    ebv = np.repeat(eb, N)
    u = np.linalg.lstsq(K, -Dv @ C @ ebv, rcond=1e-12)[0]
    return (ebv + B @ u).reshape((6,) + gs.shape)


@pytest.mark.parametrize("dims", [(4, 4, 4), (5, 5, 5), (6, 4, 5), (6, 6, 6)])
def test_brute_force_dense(dims):
    gs = O.GridSpec(*dims, 1.0, 1.3, 0.8)
    rng = np.random.default_rng(sum(dims))
    lm, mu = material(dims, rng, 1.0, 8.0)
    eb = rng.normal(size=6)
    ref = dense_solution(lm, mu, eb, gs)
    for depth in (0, 4):
        eps, _, it, _ = O.lippmann_schwinger_rfft(lm, mu, eb, gs, rtol=1e-14, depth=depth, maxiter=5000)
        assert np.abs(eps - ref).max() < 1e-9 * np.abs(ref).max(), (depth, it)


def laminate_ceff(lz, mz):
    """Closed form for a voxel laminate stacked along z (Backus average)."""
    C = np.zeros((6, 6))
    a = np.mean(1 / (lz + 2 * mz))
    b = np.mean(lz / (lz + 2 * mz))
    for k in range(6):
        e = np.zeros(6)
        e[k] = 1.0
        S = (e[2] + (e[0] + e[1]) * b) / a
        ezz = (S - lz * (e[0] + e[1])) / (lz + 2 * mz)
        tr = e[0] + e[1] + ezz
        s = np.zeros(6)
        s[0] = np.mean(lz * tr + 2 * mz * e[0])
        s[1] = np.mean(lz * tr + 2 * mz * e[1])
        s[2] = S
        s[3] = 2 * np.mean(mz) * e[3]
        s[4] = 2 * e[4] / np.mean(1 / mz)
        s[5] = 2 * e[5] / np.mean(1 / mz)
        C[k] = s
    return C


def test_laminate_closed_form():
    gs = O.GridSpec(6, 5, 16, 1.0, 1.0, 2.0)
    rng = np.random.default_rng(4)
    lz = rng.uniform(1.0, 20.0, size=gs.nz)
    mz = rng.uniform(1.0, 30.0, size=gs.nz)
    lm = np.broadcast_to(lz, gs.shape).copy()
    mu = np.broadcast_to(mz, gs.shape).copy()
    C, its = O.effective_tensor(lm, mu, gs, rtol=1e-13, maxiter=5000)
    Cref = laminate_ceff(lz, mz)
    assert np.abs(C - Cref).max() < 1e-9 * np.abs(Cref).max()
    # tensor-shear Voigt: symmetric after weighting the shear columns (Hill-Mandel)
    Ceng = C.T @ np.diag(1.0 / O.VOIGT_WEIGHTS)
    assert np.allclose(Ceng, Ceng.T, atol=1e-9 * np.abs(Cref).max())


def test_voigt_reuss_bounds_and_symmetry():
    gs = O.GridSpec(10, 9, 8, 1.0, 1.0, 1.0)
    rng = np.random.default_rng(9)
    lm, mu = material(gs.shape, rng, 1.0, 10.0)
    C, _ = O.effective_tensor(lm, mu, gs, rtol=1e-12, maxiter=5000)
    # sigma = C^T eps_tensor = C^T diag(1/w) eps_engineering: the engineering-shear
    # stiffness C^T diag(1/w) is symmetric by Hill-Mandel
    Cs = C.T @ np.diag(1.0 / O.VOIGT_WEIGHTS)
    assert np.allclose(Cs, Cs.T, atol=1e-9 * np.abs(Cs).max())
    Cv = np.zeros((6, 6))
    Sr = np.zeros((6, 6))
    for l, m in zip(lm.ravel(), mu.ravel()):
        c = np.zeros((6, 6))
        c[:3, :3] = l
        c[np.arange(3), np.arange(3)] += 2 * m
        c[np.arange(3, 6), np.arange(3, 6)] = m  # engineering shear stiffness
        Cv += c
        Sr += np.linalg.inv(c)
    Cv /= lm.size
    Cr = np.linalg.inv(Sr / lm.size)
    ev_up = np.linalg.eigvalsh(Cv - Cs)
    ev_lo = np.linalg.eigvalsh(Cs - Cr)
    assert ev_up.min() > -1e-9 and ev_lo.min() > -1e-9


# --------------------------------------------------------------------------- Anderson
def test_anderson_depth0_is_plain():
    gs = O.GridSpec(10, 8, 6, 1.0, 1.0, 1.0)
    rng = np.random.default_rng(2)
    lm, mu = material(gs.shape, rng, 1.0, 5.0)
    eb = rng.normal(size=6)
    e1, s1, i1 = O.lippmann_schwinger_capi(lm, mu, eb, gs, 1e-8, 0.0, 500)
    e2, s2, i2, _ = O.lippmann_schwinger_rfft(lm, mu, eb, gs, rtol=1e-8, depth=0, maxiter=500)
    assert i1 == i2
    assert np.array_equal(e1, e2)


def test_anderson_full_memory_equals_gmres():
    """Walker & Ni (2011): with full memory on a linear fixed-point map g, the AA iterate
    x_{k+1} equals g(x_k^GMRES), GMRES minimising the same (weighted) residual norm."""
    gs = O.GridSpec(4, 5, 4, 1.0, 1.0, 1.0)
    rng = np.random.default_rng(11)
    lm, mu = material(gs.shape, rng, 1.0, 6.0)
    eb = rng.normal(size=6)
    lam0, mu0 = 0.5 * (lm.min() + lm.max()), 0.5 * (mu.min() + mu.max())
    xz = O.get_xizero(gs)
    w = np.sqrt(O.VOIGT_WEIGHTS)[:, None, None, None]

    def resid(x):  # r(x) = -Gamma0 C x
        return np.real(np.fft.ifftn(O.fourier_solve(np.fft.fftn(O.compute_sigma(lm, mu, x.copy()), axes=(1, 2, 3)),
                                                    lam0, mu0, xz), axes=(1, 2, 3)))

    x0 = np.broadcast_to(eb[:, None, None, None], (6,) + gs.shape).copy()
    r0 = resid(x0)
    A = lambda d: -(resid(x0 + d) - r0)  # noqa: E731  linear part
    for k in range(1, 5):
        K = [r0]
        for _ in range(k - 1):
            K.append(A(K[-1]))
        AK = np.stack([(w * A(v)).ravel() for v in K], axis=1)
        c = np.linalg.lstsq(AK, (w * r0).ravel(), rcond=None)[0]
        xg = x0 + sum(ci * v for ci, v in zip(c, K))
        g_xg = xg + resid(xg)
        eaa, _, it, _ = O.lippmann_schwinger_rfft(lm, mu, eb, gs, rtol=0.0, atol=0.0, depth=8, maxiter=k + 1)
        assert it == k + 1
        assert np.abs(eaa - g_xg).max() < 1e-10 * np.abs(g_xg).max()


# --------------------------------------------------------------------------- rasteriser
def test_rasteriser_matches_reference_material_properties():
    """The generator's rasteriser against the literal restatement of data.py:267-314."""
    from oracle import microstructure_oracle as MO

    for shape, nl, r, vf, seed in [((32, 32, 32), 4, 2.5, 0.4, 1), ((40, 36, 48), 3, 4.0, 0.5, 2)]:
        ext = tuple(float(n) for n in shape)
        lam, mu, vfa, fib = MO.crossply(shape, ext, nl, r, vf, seed)
        pos = [[] for _ in range(nl)]
        for (k, c, z) in fib:
            pos[k].append((c, z))
        gs = O.GridSpec(*shape, *ext)
        ref = O.material_properties(gs, [np.array(p).reshape(-1, 2) for p in pos],
                                    [np.full(len(p), r) for p in pos], [k % 2 for k in range(nl)],
                                    values_fibre=(29.17,), values_material=(1.111,))
        assert np.array_equal(ref[0] != 1.111, mu != np.float32(1.111))
        assert abs(vfa - (mu != np.float32(1.111)).mean()) < 1e-15


def test_rsa_invariants():
    from oracle import microstructure_oracle as MO

    shape, ext, nl, r, vf = (64, 64, 64), (64.0, 64.0, 64.0), 4, 3.0, 0.5
    fib = MO.place(shape, ext, nl, r, vf, 123)
    t = ext[2] / nl
    for k in range(nl):
        f = [(c, z) for (kk, c, z) in fib if kk == k]
        Lc = ext[1] if k % 2 else ext[0]
        assert len(f) <= round(vf * Lc * t / (np.pi * r * r))
        for i, (c, z) in enumerate(f):
            assert k * t + r - 1e-12 <= z <= (k + 1) * t - r + 1e-12
            for (c2, z2) in f[i + 1:]:
                dc = abs(c - c2)
                dc = min(dc, Lc - dc)
                assert dc * dc + (z - z2) ** 2 >= 4 * r * r


# ------------------------------------------------------------------ C_eff gradient
def test_ceff_gradient_closed_form_matches_finite_differences():
    """ceff_gradient_closed_form against central differences of the oracle's
    C_eff (FP64, relative step 1e-4) on a 6x5x4 grid with a random two-phase
    medium, for a loss sum(W * C_eff) with every entry of W nonzero (so a
    dropped Voigt weight, a missing factor 2 or a transposed C_eff index
    changes the answer)."""
    gs = O.GridSpec(6, 5, 4, 6.0, 5.0, 4.0)
    rng = np.random.default_rng(42)
    phase = rng.random(gs.shape) < 0.4
    lam = np.where(phase, 19.44, 2.593) * rng.uniform(0.9, 1.1, gs.shape)
    mu = np.where(phase, 29.17, 1.111) * rng.uniform(0.9, 1.1, gs.shape)
    W = rng.standard_normal((6, 6))

    def ceff(lm, m):
        return O.effective_tensor(lm, m, gs, rtol=1e-13, depth=4, maxiter=5000)[0]

    eps = np.zeros((6, 6) + gs.shape)
    for k in range(6):
        e = np.zeros(6)
        e[k] = 1.0
        eps[k] = O.lippmann_schwinger_rfft(lam, mu, e, gs, rtol=1e-13, depth=4, maxiter=5000)[0]
    g_lam, g_mu = O.ceff_gradient_closed_form(eps, W)
    voxels = [(0, 0, 0), (3, 2, 1), (5, 4, 3), (2, 1, 2)]
    for field, grad in (("lam", g_lam), ("mu", g_mu)):
        for v in voxels:
            lp, lm_, mp, mm = lam.copy(), lam.copy(), mu.copy(), mu.copy()
            base = lam[v] if field == "lam" else mu[v]
            h = 1e-4 * base
            if field == "lam":
                lp[v] += h
                lm_[v] -= h
            else:
                mp[v] += h
                mm[v] -= h
            fd = float((W * (ceff(lp, mp) - ceff(lm_, mm))).sum()) / (2 * h)
            assert abs(grad[v] - fd) <= 1e-6 * max(abs(fd), 1e-3), (field, v, grad[v], fd)


# ------------------------------------------------------------------ conjugate gradients
def _energy(x, lm, mu):
    return float(np.einsum("aijk,aijk,a->", x, O.compute_sigma(lm, mu, x.copy()), O.VOIGT_WEIGHTS))


@pytest.mark.parametrize("dims", [(4, 4, 4), (5, 4, 3)])
def test_cg_iterates_are_krylov_galerkin_optimal(dims):
    """The k-th CG iterate minimises the energy norm of the error
    <e - e*, C (e - e*)>_w over eps_bar + span{r0, A r0, ..., A^(k-1) r0},
    A = Gamma0 C (the defining property of CG in the C0 product, in which A is
    self-adjoint): the minimiser is found by a dense least-squares solve over
    an explicit Krylov basis, independent of CG's recurrences (a wrong alpha
    or beta, or the plain product in place of C0, breaks it); e* comes from
    the dense brute-force solve of the discrete operator."""
    gs = O.GridSpec(*dims, 1.0, 1.3, 0.8)
    rng = np.random.default_rng(sum(dims) + 11)
    lm, mu = material(dims, rng, 1.0, 8.0)
    eb = rng.normal(size=6)
    estar = dense_solution(lm, mu, eb, gs) - eb[:, None, None, None]
    sp = O._RealSpectral(gs)
    l0, m0 = 0.5 * (lm.min() + lm.max()), 0.5 * (mu.min() + mu.max())
    update = lambda s: sp.inv(O.fourier_solve(sp.fwd(s), l0, m0, sp.xizero))  # noqa: E731
    ebf = np.broadcast_to(eb[:, None, None, None], (6,) + gs.shape).copy()
    r0 = update(O.compute_sigma(lm, mu, ebf.copy()))
    basis = [r0]
    for k in range(1, 5):
        eps_k = O.lippmann_schwinger_cg(lm, mu, eb, gs, rtol=0.0, maxiter=k)[0] - ebf
        # energy-optimal combination of the basis (Gram system in the energy product)
        G = np.array([[_energy(a + b, lm, mu) - _energy(a - b, lm, mu) for b in basis] for a in basis]) / 4
        rhs = np.array([(_energy(a + estar, lm, mu) - _energy(a - estar, lm, mu)) / 4 for a in basis])
        y = np.linalg.solve(G, rhs)
        best = sum(c * v for c, v in zip(y, basis))
        assert np.abs(eps_k - best).max() <= 1e-8 * np.abs(estar).max(), k
        basis.append(-update(O.compute_sigma(lm, mu, basis[-1].copy())))  # A^k r0


def test_cg_converges_to_the_plain_solution_and_counts_updates():
    gs = py_grid("odd")
    rng = np.random.default_rng(784173)
    lm, mu = material(gs.shape, rng)
    eb = np.array([1.0, 0.4, 0.3, 0.1, -0.4, 0.7])
    e_cg, s_cg, it, hist = O.lippmann_schwinger_cg(lm, mu, eb, gs, rtol=1e-12, maxiter=200)
    e_pl, _, _, _ = O.lippmann_schwinger_rfft(lm, mu, eb, gs, rtol=1e-12, depth=0, maxiter=2000)
    assert np.abs(e_cg - e_pl).max() <= 1e-9 * np.abs(e_pl).max()
    assert len(hist) == it + 1 and hist[-1] <= 1e-12 * hist[0] < hist[-2]
    # faster than the basic scheme on the reference test medium
    _, _, it_pl, _ = O.lippmann_schwinger_rfft(lm, mu, eb, gs, rtol=1e-6, depth=0, maxiter=2000)
    assert O.lippmann_schwinger_cg(lm, mu, eb, gs, rtol=1e-6, maxiter=200)[2] < it_pl
