"""FP64 CPU oracle for the Moulinec-Suquet Lippmann-Schwinger solver.

TEST INFRASTRUCTURE ONLY.  Nothing in the product path (the HIP library, the
Python loader in ``jaxmaterials_amd``) imports or calls this module.  It is
used by ``tests/`` as the checker, by ``__graft_entry__.smoke()`` to check the
GPU result, and by ``bench.py`` for the ``cpu_baseline`` leg only.

Every function restates the reference algorithm in float64 NumPy, citing the
reference file and lines it follows (paths relative to the reference repo):

* ``compute_sigma``              src/jaxmaterials/solver/lippmann_schwinger.py:20-33
* ``backward_derivative`` /
  ``backward_divergence``        src/jaxmaterials/solver/derivatives.py:8-96
* ``relative_divergence``        src/jaxmaterials/solver/lippmann_schwinger.py:103-115
* ``relative_divergence_fourier`` src/jaxmaterials/solver/lippmann_schwinger.py:118-143
* ``get_xi`` / ``get_xizero``    src/jaxmaterials/solver/fourier.py:9-85
* ``fourier_solve``              src/jaxmaterials/solver/fourier.py:115-186
* ``lippmann_schwinger_jax``     src/jaxmaterials/solver/lippmann_schwinger.py:146-265
* ``lippmann_schwinger_capi``    cuda/src/lippmann_schwinger.cu:166-234 (loop and
                                 return semantics of the C ABI)
* ``lippmann_schwinger_cg``      conjugate gradients on the same equation (no reference
                                 counterpart; Krylov-Galerkin optimality pinned on a dense grid)
* ``effective_tensor``           src/jaxmaterials/data/data.py:178-186
* ``material_properties``        src/jaxmaterials/data/data.py:267-314

Deliberate deviation (one): the zero set of xi-tilde is decided exactly on the
integer mode indices (k = 0, or at least two axes at their Nyquist index)
instead of the float threshold of fourier.py:84 / fourier.cu:68-72.  The
threshold form is kept as ``get_xizero_thresholded`` and a test shows the two
agree to 1e-12 on the reference's test grids.

Pins (what checks each function; all in tests/test_oracle.py):

* Hooke: closed form on random inputs (isotropic C as an explicit 6x6 matrix).
* xi / xizero: closed form, exact zero set, thresholded reference form.
* fourier_solve (Gamma0): an independent tensor-index construction of
  Gamma0_{klij} (doc/linear_elasticity.tex:137-152) contracted with einsum;
  the projector identity Gamma0 C0 Gamma0 = Gamma0; TestDivSigma's invariant
  (cuda/src/tests/test_fourier.hh:267-316).
* residual: real-space vs Fourier (tests/test_lippmann_schwinger.py:50-61).
* whole solve: homogeneous material converges in 0 iterations
  (cuda/src/tests/test_lippmann_schwinger.hh:104-128); a dense brute-force solve
  of the discrete D+/D- problem on 4^3..6^3 grids; the voxel laminate closed
  form; iteration bounds of tests/test_lippmann_schwinger.py:64-99.
* Anderson: depth 0 is identical to plain; with full memory on a linear map
  the iterates equal g(x_k^GMRES) (Walker & Ni 2011).
* rasteriser: brute-force per-voxel loop on small grids.
"""

import numpy as np

__all__ = [
    "GridSpec",
    "VOIGT_WEIGHTS",
    "compute_sigma",
    "backward_derivative",
    "backward_divergence",
    "relative_divergence",
    "relative_divergence_fourier",
    "relative_divergence_rfft",
    "axis_tables",
    "get_xi",
    "get_xizero",
    "get_xizero_thresholded",
    "fourier_solve",
    "lippmann_schwinger_jax",
    "lippmann_schwinger_capi",
    "lippmann_schwinger_rfft",
    "effective_tensor",
    "material_properties",
]

# Inner products and norms of Voigt vectors with tensor shear components weight
# the three shear entries by 2 (lippmann_schwinger.py:113, 141, 220; common.cc:36-58).
VOIGT_WEIGHTS = np.array([1.0, 1.0, 1.0, 2.0, 2.0, 2.0])


class GridSpec:
    """Structured periodic grid (src/jaxmaterials/common.py:6-29)."""

    def __init__(self, nx, ny, nz, Lx, Ly, Lz):
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.Lx, self.Ly, self.Lz = float(Lx), float(Ly), float(Lz)

    @property
    def number_of_voxels(self):
        return self.nx * self.ny * self.nz

    @property
    def shape(self):
        return (self.nx, self.ny, self.nz)


def _shape(gs):
    """(nx, ny, nz) of any grid-spec object with nx, ny, nz attributes."""
    return (gs.nx, gs.ny, gs.nz)


def compute_sigma(lmbda, mu, epsilon):
    """sigma_i = 2 mu eps_i + lambda tr(eps) [i<3]  (lippmann_schwinger.py:20-33)."""
    tr = epsilon[0] + epsilon[1] + epsilon[2]
    sigma = 2.0 * mu * epsilon
    sigma[:3] += lmbda * tr
    return sigma


def backward_derivative(g, grid_spec, direction):
    """Backward derivative D^-_d on the rotated staggered grid (derivatives.py:8-63).

    D^-_0 g = (n_x/(4 L_x)) (1 + S_y)(1 + S_z)(1 - S_x) g where S_d shifts by +1
    (np.roll(g, 1, axis)).
    """
    ax = [-3, -2, -1]
    n = [grid_spec.nx, grid_spec.ny, grid_spec.nz][direction]
    L = [grid_spec.Lx, grid_spec.Ly, grid_spec.Lz][direction]
    others = [a for i, a in enumerate(ax) if i != direction]
    d = ax[direction]
    s = g + np.roll(g, 1, axis=others[0])
    s = s + np.roll(s, 1, axis=others[1])
    return 0.25 * n / L * (s - np.roll(s, 1, axis=d))


def backward_divergence(sigma, grid_spec):
    """Row-wise D^- divergence of a Voigt tensor field (derivatives.py:66-96)."""
    D = lambda f, d: backward_derivative(f, grid_spec, d)  # noqa: E731
    return np.stack(
        [
            D(sigma[0], 0) + D(sigma[3], 1) + D(sigma[4], 2),
            D(sigma[3], 0) + D(sigma[1], 1) + D(sigma[5], 2),
            D(sigma[4], 0) + D(sigma[5], 1) + D(sigma[2], 2),
        ]
    )


def relative_divergence(sigma, grid_spec):
    """||D^- . sigma|| / (sqrt(N) ||<sigma>||_V)  (lippmann_schwinger.py:103-115)."""
    dsig = backward_divergence(sigma, grid_spec)
    num = np.sqrt(np.sum(dsig**2))
    avg = np.mean(sigma, axis=(1, 2, 3))
    den = np.sqrt(np.sum(VOIGT_WEIGHTS * avg**2))
    return num / (np.sqrt(grid_spec.number_of_voxels) * den)


def axis_tables(n, L):
    """Per-axis factors of xi-tilde with exact zeros.

    Returns (g, s, c) with g = 2 n / L, s[k] = sin(pi k / n), c[k] = cos(pi k / n),
    s[0] = 0 and c[n/2] = 0 exactly (n even).  xi-tilde_d(k) = g_d s_d c_e c_f
    (fourier.py:9-58, doc/linear_elasticity.tex:97-108).
    """
    k = np.arange(n)
    s = np.sin(np.pi * k / n)
    c = np.cos(np.pi * k / n)
    s[0] = 0.0
    if n % 2 == 0:
        c[n // 2] = 0.0
    return 2.0 * n / L, s, c


def get_xi(grid_spec, kz_half=False):
    """xi-tilde on the full (or z-half) spectrum, shape (3, nx, ny, nz[/2+1]).

    Follows fourier.py:9-58 (the same closed form) with exact zeros.
    """
    gx, sx, cx = axis_tables(grid_spec.nx, grid_spec.Lx)
    gy, sy, cy = axis_tables(grid_spec.ny, grid_spec.Ly)
    gz, sz, cz = axis_tables(grid_spec.nz, grid_spec.Lz)
    if kz_half:
        m = grid_spec.nz // 2 + 1
        sz, cz = sz[:m], cz[:m]
    SX, SY, SZ = np.meshgrid(sx, sy, sz, indexing="ij")
    CX, CY, CZ = np.meshgrid(cx, cy, cz, indexing="ij")
    return np.stack([gx * SX * CY * CZ, gy * CX * SY * CZ, gz * CX * CY * SZ])


def get_xizero(grid_spec, kz_half=False):
    """Normalised xi-tilde with the exact zero set (fourier.py:61-85, tex:148-152)."""
    xi = get_xi(grid_spec, kz_half=kz_half)
    nrm = np.sqrt(np.sum(xi**2, axis=0))
    out = np.zeros_like(xi)
    nz = nrm > 0.0
    out[:, nz] = xi[:, nz] / nrm[nz]
    return out


def get_xizero_thresholded(grid_spec):
    """Literal restatement of fourier.py:9-85 (np.sin/np.cos, threshold 1e-12)."""
    K = [2 * np.pi * np.arange(n) / n for n in _shape(grid_spec)]
    X = np.meshgrid(*K, indexing="ij")
    xi = np.stack(
        [
            2 * grid_spec.nx / grid_spec.Lx * np.sin(X[0] / 2) * np.cos(X[1] / 2) * np.cos(X[2] / 2),
            2 * grid_spec.ny / grid_spec.Ly * np.cos(X[0] / 2) * np.sin(X[1] / 2) * np.cos(X[2] / 2),
            2 * grid_spec.nz / grid_spec.Lz * np.cos(X[0] / 2) * np.cos(X[1] / 2) * np.sin(X[2] / 2),
        ]
    )
    nrm = np.linalg.norm(xi, axis=0)
    nrm[nrm < 1.0e-12] = 1
    return xi / nrm


def relative_divergence_fourier(sigma_hat, xi, grid_spec=None):
    """Full-spectrum residual (lippmann_schwinger.py:118-143)."""
    d0 = xi[0] * sigma_hat[0] + xi[1] * sigma_hat[3] + xi[2] * sigma_hat[4]
    d1 = xi[0] * sigma_hat[3] + xi[1] * sigma_hat[1] + xi[2] * sigma_hat[5]
    d2 = xi[0] * sigma_hat[4] + xi[1] * sigma_hat[5] + xi[2] * sigma_hat[2]
    num = np.sqrt(np.sum(np.abs(d0) ** 2 + np.abs(d1) ** 2 + np.abs(d2) ** 2))
    s0 = np.real(sigma_hat[:, 0, 0, 0])
    return num / np.sqrt(np.sum(VOIGT_WEIGHTS * s0**2))


def hermitian_weights(nz):
    """Weights of the z-half spectrum: 1 for kz = 0 and (nz even) kz = nz/2, else 2.

    Same rule as reduce_fourier_kernel (cuda/src/fourier.cu:155-159).
    """
    m = nz // 2 + 1
    w = np.full(m, 2.0)
    w[0] = 1.0
    if nz % 2 == 0:
        w[nz // 2] = 1.0
    return w


def relative_divergence_rfft(sigma_hat_half, xi_half, nz):
    """Residual on the z-half spectrum with Hermitian weights.

    Equal to relative_divergence_fourier on the full spectrum (the CUDA path's
    divergence_fourier_kernel + reduce_fourier, fourier.cu:140-265).  Returns
    (R, ||sigma_hat(0)||_V).
    """
    d0 = xi_half[0] * sigma_hat_half[0] + xi_half[1] * sigma_hat_half[3] + xi_half[2] * sigma_hat_half[4]
    d1 = xi_half[0] * sigma_hat_half[3] + xi_half[1] * sigma_hat_half[1] + xi_half[2] * sigma_hat_half[5]
    d2 = xi_half[0] * sigma_hat_half[4] + xi_half[1] * sigma_hat_half[5] + xi_half[2] * sigma_hat_half[2]
    e = np.abs(d0) ** 2 + np.abs(d1) ** 2 + np.abs(d2) ** 2
    num = np.sqrt(np.sum(e * hermitian_weights(nz)))
    s0 = np.real(sigma_hat_half[:, 0, 0, 0])
    den = np.sqrt(np.sum(VOIGT_WEIGHTS * s0**2))
    # a zero load gives 0 / 0 = nan, as in the reference; its loop then does
    # not run
    with np.errstate(invalid="ignore", divide="ignore"):
        return num / den, den


def fourier_solve(tau_hat, lmbda0, mu0, xizero):
    """r-hat = -Gamma0 tau-hat for the isotropic reference medium (fourier.py:115-186).

    Component form of doc/linear_elasticity.tex:205-226:
    (1/mu0) (-A(xi0) tau + (lambda0+mu0)/(lambda0+2mu0) B(xi0) tau).
    """
    x0, x1, x2 = xizero[0], xizero[1], xizero[2]
    t = tau_hat
    eA = np.stack(
        [
            x0**2 * t[0] + x0 * (x2 * t[4] + x1 * t[3]),
            x1**2 * t[1] + x1 * (x2 * t[5] + x0 * t[3]),
            x2**2 * t[2] + x2 * (x1 * t[5] + x0 * t[4]),
            0.5 * (x0 * x1 * (t[0] + t[1]) + (x0**2 + x1**2) * t[3] + x2 * (x0 * t[5] + x1 * t[4])),
            0.5 * (x0 * x2 * (t[0] + t[2]) + (x0**2 + x2**2) * t[4] + x1 * (x0 * t[5] + x2 * t[3])),
            0.5 * (x1 * x2 * (t[1] + t[2]) + (x1**2 + x2**2) * t[5] + x0 * (x1 * t[4] + x2 * t[3])),
        ]
    )
    Xi = np.stack([x0**2, x1**2, x2**2, x0 * x1, x0 * x2, x1 * x2])
    rho = x0**2 * t[0] + x1**2 * t[1] + x2**2 * t[2] + 2 * (x0 * x1 * t[3] + x0 * x2 * t[4] + x1 * x2 * t[5])
    return (1.0 / mu0) * (-eA + (lmbda0 + mu0) / (lmbda0 + 2 * mu0) * Xi * rho)


def _reference_medium(lmbda, mu):
    """lambda0, mu0 = mid-range of the fields (lippmann_schwinger.py:171-172)."""
    return 0.5 * (np.min(lmbda) + np.max(lmbda)), 0.5 * (np.min(mu) + np.max(mu))


def lippmann_schwinger_jax(lmbda, mu, epsilon_bar, grid_spec, rtol=1e-6, atol=1e-20,
                           depth=0, maxiter=32):
    """Float64 restatement of lippmann_schwinger_jax (lippmann_schwinger.py:146-265).

    Full complex FFTs (jnp.fft.fftn / ifftn) as in the reference; small grids
    only.  Returns (epsilon, sigma, iter, rel_error).
    """
    lmbda = np.asarray(lmbda, dtype=np.float64)
    mu = np.asarray(mu, dtype=np.float64)
    lmbda0, mu0 = _reference_medium(lmbda, mu)
    mu0, lmbda0 = 0.5 * (mu.min() + mu.max()), 0.5 * (lmbda.min() + lmbda.max())
    xizero = get_xizero(grid_spec)
    xi = get_xi(grid_spec)
    shp = (depth + 1, 6) + _shape(grid_spec)
    epsilon = np.zeros(shp)
    epsilon[0] = np.asarray(epsilon_bar, dtype=np.float64)[:, None, None, None]
    residual = np.zeros(shp)
    A = np.eye(depth + 1)
    u = np.zeros(depth + 1)
    sigma = compute_sigma(lmbda, mu, epsilon[0].copy())
    sigma_hat = np.fft.fftn(sigma, axes=(-3, -2, -1))
    rel_error = relative_divergence_fourier(sigma_hat, xi)
    rel_error_0 = rel_error
    it = 0
    while (rel_error > atol) and (rel_error > rtol * rel_error_0) and (it < maxiter):
        r_hat = fourier_solve(sigma_hat, lmbda0, mu0, xizero)
        r = np.real(np.fft.ifftn(r_hat, axes=(-3, -2, -1)))
        residual = np.roll(residual, 1, axis=0)
        residual[0] = r
        A = np.roll(A, (1, 1), axis=(0, 1))
        A[0, :] = np.einsum("aijk,saijk,a->s", r, residual, VOIGT_WEIGHTS)
        A[:, 0] = A[0, :]
        u = np.roll(u, 1)
        u[0] = 1.0
        v = np.linalg.solve(A, u)
        alpha = v / np.dot(v, u)
        eps_tilde = np.einsum("s,saijk->aijk", alpha, epsilon + residual)
        epsilon = np.roll(epsilon, 1, axis=0)
        epsilon[0] = eps_tilde
        sigma = compute_sigma(lmbda, mu, epsilon[0].copy())
        sigma_hat = np.fft.fftn(sigma, axes=(-3, -2, -1))
        rel_error = relative_divergence_fourier(sigma_hat, xi)
        it += 1
    return epsilon[0], sigma, it, rel_error


class _RealSpectral:
    """Half-spectrum FFT helpers shared by the rfft restatements."""

    def __init__(self, grid_spec, workers=None):
        import scipy.fft as sfft

        self.sfft = sfft
        self.gs = grid_spec
        self.workers = workers
        self.xi = get_xi(grid_spec, kz_half=True)
        self.xizero = get_xizero(grid_spec, kz_half=True)

    def fwd(self, f):
        return self.sfft.rfftn(f, axes=(-3, -2, -1), workers=self.workers)

    def inv(self, fh):
        return self.sfft.irfftn(fh, s=_shape(self.gs), axes=(-3, -2, -1), workers=self.workers)

    def residual(self, sigma_hat):
        return relative_divergence_rfft(sigma_hat, self.xi, self.gs.nz)


def lippmann_schwinger_rfft(lmbda, mu, epsilon_bar, grid_spec, rtol=1e-6, atol=0.0,
                            depth=0, maxiter=1000, workers=None, tau=None,
                            epsilon0=None, rel_error_ref=None, max_iterations_timed=None,
                            callback=None, stress_law=None):
    """Same iteration as lippmann_schwinger_jax on real FFTs (scipy.fft, float64).

    The z-half spectrum and the Hermitian weights of fourier.cu:155-159 replace
    the full complex spectrum; the arithmetic is otherwise identical.
    Extensions used by the GPU path's parity tests (each defaults to the
    reference behaviour):

    * ``tau``: an additional polarisation field added to the stress
      (sigma = C eps + tau); used by the adjoint solve.
    * ``epsilon0``: a start field instead of the uniform epsilon_bar.
    * ``rel_error_ref``: normaliser of the relative tolerance.  The reference
      uses the residual of the start iterate (lippmann_schwinger.py:193-194);
      with a warm start the zero-start residual R(epsilon_bar) is passed here.
    * ``max_iterations_timed``: stop after that many updates (CPU timing).
    * ``stress_law``: sigma(eps) in place of Hooke's law (the damaged law of
      elasticity_solve, lippmann_schwinger.py:334-462; the reference medium
      stays the mid-range of lmbda and mu).

    Returns (epsilon, sigma, iter, history) with history the list of R values.
    """
    sp = _RealSpectral(grid_spec, workers)
    lmbda = np.asarray(lmbda, dtype=np.float64)
    mu = np.asarray(mu, dtype=np.float64)
    mu0, lmbda0 = 0.5 * (mu.min() + mu.max()), 0.5 * (lmbda.min() + lmbda.max())
    shp = (depth + 1, 6) + _shape(grid_spec)
    epsilon = np.zeros(shp)
    if epsilon0 is None:
        epsilon[0] = np.asarray(epsilon_bar, dtype=np.float64)[:, None, None, None]
    else:
        epsilon[0] = epsilon0
    residual = np.zeros(shp)
    A = np.eye(depth + 1)
    u = np.zeros(depth + 1)

    def stress(e):
        s = compute_sigma(lmbda, mu, e.copy()) if stress_law is None else stress_law(e)
        if tau is not None:
            s += tau
        return s

    sigma = stress(epsilon[0])
    sigma_hat = sp.fwd(sigma)
    rel_error, _ = sp.residual(sigma_hat)
    rel_error_0 = rel_error if rel_error_ref is None else rel_error_ref
    history = [rel_error]
    it = 0
    while (rel_error > atol) and (rel_error > rtol * rel_error_0) and (it < maxiter):
        if max_iterations_timed is not None and it >= max_iterations_timed:
            break
        r = sp.inv(fourier_solve(sigma_hat, lmbda0, mu0, sp.xizero))
        residual = np.roll(residual, 1, axis=0)
        residual[0] = r
        A = np.roll(A, (1, 1), axis=(0, 1))
        A[0, :] = np.einsum("aijk,saijk,a->s", r, residual, VOIGT_WEIGHTS)
        A[:, 0] = A[0, :]
        u = np.roll(u, 1)
        u[0] = 1.0
        v = np.linalg.solve(A, u)
        alpha = v / np.dot(v, u)
        eps_tilde = np.einsum("s,saijk->aijk", alpha, epsilon + residual)
        epsilon = np.roll(epsilon, 1, axis=0)
        epsilon[0] = eps_tilde
        sigma = stress(epsilon[0])
        sigma_hat = sp.fwd(sigma)
        rel_error, _ = sp.residual(sigma_hat)
        history.append(rel_error)
        it += 1
        if callback is not None:
            callback(it, rel_error)
    return epsilon[0], sigma, it, history


def lippmann_schwinger_capi(lmbda, mu, epsilon_bar, grid_spec, rtol, atol, maxiter):
    """Loop and return semantics of the C ABI (cuda/src/lippmann_schwinger.cu:166-234).

    for iter < maxiter: sigma = C eps; sigma_hat = FFT; R = residual;
    R0 = R at iter 0; break if R < max(rtol R0, atol); eps += IFFT(-Gamma0 sigma_hat).
    Returns (epsilon, sigma, iter).  When the loop runs out, sigma is the stress
    of the iterate before the last update (the reference never recomputes it).
    Float64, real FFTs.
    """
    sp = _RealSpectral(grid_spec)
    lmbda = np.asarray(lmbda, dtype=np.float64)
    mu = np.asarray(mu, dtype=np.float64)
    # lippmann_schwinger.cu:172-173
    lmbda0 = 0.5 * (lmbda.max() + lmbda.min())
    mu0 = 0.5 * (mu.max() + mu.min())
    eps = np.empty((6,) + _shape(grid_spec))
    eps[:] = np.asarray(epsilon_bar, dtype=np.float64)[:, None, None, None]
    sigma = None
    R0 = None
    it = 0
    for it in range(maxiter):
        sigma = compute_sigma(lmbda, mu, eps.copy())
        sigma_hat = sp.fwd(sigma)
        R, _ = sp.residual(sigma_hat)
        if it == 0:
            R0 = R
        if R < max(rtol * R0, atol):
            return eps, sigma, it
        eps += sp.inv(fourier_solve(sigma_hat, lmbda0, mu0, sp.xizero))
    else:
        it = maxiter
    return eps, sigma, it


def c0_dot(a, b, lmbda0, mu0):
    """<a, C0 b>_w = sum_x 2 mu0 sum_i w_i a_i b_i + lmbda0 tr a tr b (isotropic reference medium)."""
    return float(2.0 * mu0 * np.einsum("aijk,aijk,a->", a, b, VOIGT_WEIGHTS)
                 + lmbda0 * np.sum((a[0] + a[1] + a[2]) * (b[0] + b[1] + b[2])))


def lippmann_schwinger_cg(lmbda, mu, epsilon_bar, grid_spec, rtol=1e-6, atol=0.0, maxiter=1000, workers=None,
                          epsilon0=None, rel_error_ref=None, callback=None):
    """Conjugate gradients on the Lippmann-Schwinger equation (no counterpart in
    the reference, which iterates the basic scheme or Anderson mixing,
    lippmann_schwinger.py:146-265): textbook CG (Hestenes & Stiefel 1952) for
    A e = b on compatible zero-mean strains, A = Gamma0 C, b = -Gamma0 C eps_bar,
    in the reference-medium inner product <a, b>_C0 (in which A is self-adjoint
    and positive definite; Gamma0 C0 is the C0-orthogonal projection onto
    compatible fields).  The residual b - A e of eps = eps_bar + e is the
    Lippmann-Schwinger update -Gamma0 (C eps) (fourier_solve).  Same reference
    medium, residual R of the iterate's stress, stop test and update count as
    lippmann_schwinger_rfft; float64, real FFTs.  Returns (epsilon, sigma, iter,
    history of R).
    """
    sp = _RealSpectral(grid_spec, workers)
    lmbda = np.asarray(lmbda, dtype=np.float64)
    mu = np.asarray(mu, dtype=np.float64)
    mu0, lmbda0 = 0.5 * (mu.min() + mu.max()), 0.5 * (lmbda.min() + lmbda.max())
    if epsilon0 is None:
        eps = np.broadcast_to(np.asarray(epsilon_bar, dtype=np.float64)[:, None, None, None],
                              (6,) + _shape(grid_spec)).copy()
    else:
        eps = np.asarray(epsilon0, dtype=np.float64).copy()

    def update(sigma_hat):  # -Gamma0 sigma_hat in real space
        return sp.inv(fourier_solve(sigma_hat, lmbda0, mu0, sp.xizero))

    sigma = compute_sigma(lmbda, mu, eps.copy())
    sigma_hat = sp.fwd(sigma)
    R, _ = sp.residual(sigma_hat)
    R0 = R if rel_error_ref is None else rel_error_ref
    history = [R]
    r = update(sigma_hat)
    p = r.copy()
    rr = c0_dot(r, r, lmbda0, mu0)
    it = 0
    while (R > atol) and (R > rtol * R0) and (it < maxiter):
        q = -update(sp.fwd(compute_sigma(lmbda, mu, p.copy())))  # A p
        alpha = rr / c0_dot(p, q, lmbda0, mu0)
        eps += alpha * p
        r -= alpha * q
        rr_new = c0_dot(r, r, lmbda0, mu0)
        p = r + (rr_new / rr) * p
        rr = rr_new
        sigma = compute_sigma(lmbda, mu, eps.copy())
        R, _ = sp.residual(sp.fwd(sigma))
        history.append(R)
        it += 1
        if callback is not None:
            callback(it, R)
    return eps, sigma, it, history


def effective_tensor(lmbda, mu, grid_spec, rtol=1e-6, atol=0.0, depth=0, maxiter=2000,
                     workers=None):
    """6x6 C_eff: row k = <sigma> under epsilon_bar = e_k (data.py:178-186).

    Shear columns carry the factor 2 of the tensor-shear Voigt convention.
    Returns (C_eff, iterations per load case).
    """
    C = np.zeros((6, 6))
    its = []
    for k in range(6):
        e = np.zeros(6)
        e[k] = 1.0
        _, sigma, it, _ = lippmann_schwinger_rfft(lmbda, mu, e, grid_spec, rtol=rtol, atol=atol,
                                                  depth=depth, maxiter=maxiter, workers=workers)
        C[k] = sigma.mean(axis=(1, 2, 3))
        its.append(it)
    return C, its


def material_properties(grid_spec, fibre_positions, fibre_radii, fibre_orientations,
                        values_fibre, values_material, values_void=None, d_void=0.0,
                        periodic_images=True):
    """Rasterise layered fibres into per-voxel fields (data.py:267-314).

    ``values_*`` are tuples of per-channel values (e.g. (mu, lambda) as in the
    reference, data.py:295-300).  Voxel centres x = (i + 1/2) L / n; base value
    = matrix; void slabs z < d_void or z > L_z - d_void; every fibre adds
    (v_f - v_m) where (c - p0)^2 + (z - p1)^2 < r^2 (strict, additive), with c = y
    for orientation 1 and x for orientation 0.  With ``periodic_images`` the
    in-plane coordinate is also tested against the images p0 +- L (the
    reference appends boundary copies to the fibre list, data.py:239-258).
    """
    gs = grid_spec
    X, Y, Z = np.meshgrid(
        gs.Lx / gs.nx * (0.5 + np.arange(gs.nx)),
        gs.Ly / gs.ny * (0.5 + np.arange(gs.ny)),
        gs.Lz / gs.nz * (0.5 + np.arange(gs.nz)),
        indexing="ij",
    )
    nch = len(values_material)
    data = np.zeros((nch,) + _shape(gs))
    for j in range(nch):
        vm = values_material[j]
        data[j] = vm
        if values_void is not None:
            data[j] += ((Z < d_void) + (Z > gs.Lz - d_void)) * (values_void[j] - vm)
        for pos, rad, ori in zip(fibre_positions, fibre_radii, fibre_orientations):
            coord = Y if ori else X
            Lc = gs.Ly if ori else gs.Lx
            shifts = (-Lc, 0.0, Lc) if periodic_images else (0.0,)
            for p, r in zip(pos, rad):
                for sh in shifts:
                    data[j] += ((coord - p[0] - sh) ** 2 + (Z - p[1]) ** 2 < r**2) * (values_fibre[j] - vm)
    return data


def ceff_gradient_closed_form(eps_fields, dL_dC):
    """Gradient of a loss on C_eff with respect to the Lame fields, from the six
    converged strain fields alone (the build's closed form; FD-pinned in tests).

    By Hill-Mandel, W(a, b) = <eps^a : C : eps^b> = w_a C_eff[b][a] (tensor-shear
    Voigt weights w), and at equilibrium the variation of W with C(x) is
    eps^a(x) : dC(x) : eps^b(x) / N.  Hence
        dC_eff[b][a]/dlambda(x) = tr eps^a tr eps^b / (N w_a),
        dC_eff[b][a]/dmu(x)     = 2 <eps^a, eps^b>_w / (N w_a).
    eps_fields: [6, 6, nx, ny, nz] (load case, component); dL_dC: [6, 6].
    Returns (dL/dlambda, dL/dmu) as float64 arrays [nx, ny, nz].
    """
    e = np.asarray(eps_fields, dtype=np.float64)
    N = e[0, 0].size
    tr = e[:, 0] + e[:, 1] + e[:, 2]
    dl = np.zeros(e.shape[2:])
    dm = np.zeros(e.shape[2:])
    W = VOIGT_WEIGHTS
    for b in range(6):
        for a in range(6):
            c = dL_dC[b, a] / (N * W[a])
            if c == 0.0:
                continue
            dl += c * tr[a] * tr[b]
            dm += c * 2.0 * np.einsum("c,c...->...", W, e[a] * e[b])
    return dl, dm
