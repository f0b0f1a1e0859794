"""GPU parity tests of the HIP solver against the FP64 oracle.

Grids and inputs follow the reference's own tests:
* tests/test_lippmann_schwinger.py:18-47 (even [64,48,32] / odd [57,43,37],
  L = (1.2, 0.8, 0.9), lambda, mu ~ U(0.8, 1.2), seed 784173);
* cuda/src/tests/test_lippmann_schwinger.hh:19-36, 104-180 (even 48x64x32 /
  odd 47x61x37, L = (1.1, 0.9, 0.7)).
Tolerances (the bar): strain relative L2 <= 1e-5 against the FP64 oracle at the
same number of updates, iteration count within +-1, C_eff within 1e-5 relative.
"""

import numpy as np
import pytest
import torch

from jaxmaterials_amd.common import GridSpec
from jaxmaterials_amd.solver.lippmann_schwinger import lippmann_schwinger_cuda, effective_tensor
from jaxmaterials_amd.solver.device import LSSolver, unit_load_cases
from oracle import ls_oracle as O

pytestmark = pytest.mark.gpu

PY_GRIDS = {"even": (64, 48, 32), "odd": (57, 43, 37)}
CU_GRIDS = {"even": (48, 64, 32), "odd": (47, 61, 37)}
EPS_BAR = np.array([2.1, 0.9, 0.8, 0.4, 0.9, 0.5])


def py_grid(name):
    return GridSpec(*PY_GRIDS[name], 1.2, 0.8, 0.9)


def cu_grid(name):
    return GridSpec(*CU_GRIDS[name], 1.1, 0.9, 0.7)


def random_material(gs, seed=784173, lo=0.8, hi=1.2):
    """initialise_material of tests/test_lippmann_schwinger.py:37-47 (returns mu, lambda)."""
    rng = np.random.default_rng(seed=seed)
    shape = (gs.nx, gs.ny, gs.nz)
    mu = np.ones(shape) + rng.uniform(size=shape, low=lo - 1.0, high=hi - 1.0)
    lmbda = np.ones(shape) + rng.uniform(size=shape, low=lo - 1.0, high=hi - 1.0)
    return mu.astype(np.float32), lmbda.astype(np.float32)


def rel_l2(a, b):
    return float(np.sqrt(np.sum((np.asarray(a, np.float64) - b) ** 2) / np.sum(np.asarray(b, np.float64) ** 2)))


def oracle_fixed_updates(lmbda, mu, eps_bar, gs, k, depth=0, tau=None):
    """FP64 oracle strain after exactly k updates."""
    eps, sigma, it, hist = O.lippmann_schwinger_rfft(
        lmbda.astype(np.float64), mu.astype(np.float64), eps_bar, gs, rtol=0.0, atol=0.0, depth=depth,
        maxiter=k, tau=tau)
    assert it == k
    return eps, sigma, hist


@pytest.mark.parametrize("grid", ["even", "odd"])
def test_homogeneous_material_zero_iterations(gpu, grid):
    """Known answer of TestHomogeneousMaterial (cuda/src/tests/test_lippmann_schwinger.hh:104-128)."""
    gs = cu_grid(grid)
    shape = (gs.nx, gs.ny, gs.nz)
    lam = np.full(shape, 1.2, np.float32)
    mu = np.full(shape, 1.2, np.float32)
    eb = np.array([1.0, 0.4, 0.3, 0.1, -0.4, 0.7], np.float32)
    eps, sigma, it = lippmann_schwinger_cuda(lam, mu, eb, gs, rtol=1e-20, atol=1e-4)
    print(f"homogeneous {grid}: iter={it}", flush=True)
    assert it == 0
    assert np.allclose(eps, eb[:, None, None, None], rtol=0, atol=1e-7)
    s = O.compute_sigma(lam.astype(np.float64), mu.astype(np.float64), np.broadcast_to(
        eb[:, None, None, None].astype(np.float64), (6,) + shape).copy())
    assert np.allclose(sigma, s, rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("grid", ["even", "odd"])
def test_convergence_cuda_test(gpu, grid):
    """TestConvergence (cuda/src/tests/test_lippmann_schwinger.hh:132-180): iter < 10 and
    ||D- sigma|| / ||<sigma>||_V < 1e-2 for lambda, mu ~ U(0.8, 0.9)."""
    gs = cu_grid(grid)
    rng = np.random.default_rng(7812481)
    shape = (gs.nx, gs.ny, gs.nz)
    mu = rng.uniform(0.8, 0.9, size=shape).astype(np.float32)
    lam = rng.uniform(0.8, 0.9, size=shape).astype(np.float32)
    eb = np.array([1.0, 0.4, 0.3, 0.1, -0.4, 0.7], np.float32)
    eps, sigma, it = lippmann_schwinger_cuda(lam, mu, eb, gs, rtol=1e-20, atol=1e-4)
    div = O.backward_divergence(sigma.astype(np.float64), gs)
    avg = sigma.astype(np.float64).mean(axis=(1, 2, 3))
    ratio = np.sqrt(np.sum(div**2)) / np.sqrt(np.sum(O.VOIGT_WEIGHTS * avg**2))
    print(f"convergence {grid}: iter={it} ratio={ratio:.3e}", flush=True)
    assert it < 10
    assert ratio < 1e-2


@pytest.mark.parametrize("grid", ["even", "odd"])
def test_legacy_abi_matches_oracle(gpu, grid):
    """test_jax_matches_cuda (tests/test_lippmann_schwinger.py:102-141) with the FP64
    oracle in place of JAX and the bar tightened to 1e-5 relative L2 of the strain."""
    gs = py_grid(grid)
    lmbda, mu = random_material(gs)  # reference quirk Q2: the names are swapped there too
    atol = 1e-5
    eps, sigma, it = lippmann_schwinger_cuda(lmbda, mu, EPS_BAR.astype(np.float32), gs, rtol=1e-20,
                                             atol=atol, maxiter=32)
    # the loader passes mu in the lambda slot (lippmann_schwinger.py:310-312)
    e_ref, s_ref, it_ref = O.lippmann_schwinger_capi(mu, lmbda, EPS_BAR, gs, rtol=1e-20, atol=atol, maxiter=32)
    e_k, s_k, _ = oracle_fixed_updates(mu, lmbda, EPS_BAR, gs, it)
    err_e = rel_l2(eps, e_k)
    err_s = rel_l2(sigma, s_k)
    print(f"legacy {grid}: iter gpu={it} oracle={it_ref} strain relL2={err_e:.2e} stress relL2={err_s:.2e}",
          flush=True)
    assert abs(it - it_ref) <= 1
    assert err_e <= 1e-5
    assert err_s <= 1e-5


@pytest.mark.parametrize("grid", ["even", "odd"])
def test_gamma0_divergence_free(gpu, grid):
    """TestDivSigma invariant (cuda/src/tests/test_fourier.hh:267-316) through the solver:
    in a homogeneous medium with a random polarisation tau, one update gives
    sigma = C0 eps + tau divergence free."""
    gs = cu_grid(grid)
    shape = (gs.nx, gs.ny, gs.nz)
    rng = np.random.default_rng(12)
    tau = rng.normal(size=(1, 6) + shape).astype(np.float32)
    lam = torch.full(shape, 0.8, device=gpu)
    mu = torch.full(shape, 1.3, device=gpu)
    s = LSSolver(gs, nrhs=1)
    out = s.solve(lam, mu, np.zeros((1, 6)), tau=torch.from_numpy(tau).to(gpu), want_sigma=True,
                  rtol=0.0, atol=0.0, maxiter=1, check=False)
    sig = out["sigma"].cpu().numpy()[0].astype(np.float64)
    div = O.backward_divergence(sig, gs)
    rel = np.sqrt(np.sum(div**2)) / np.sqrt(np.sum(O.VOIGT_WEIGHTS[:, None, None, None] * sig**2))
    e_k, s_k, _ = oracle_fixed_updates(np.full(shape, 0.8), np.full(shape, 1.3), np.zeros(6), gs, 1,
                                       tau=tau[0].astype(np.float64))
    err = rel_l2(out["epsilon"].cpu().numpy()[0], e_k)
    print(f"divsigma {grid}: rel div={rel:.2e} strain relL2 vs oracle={err:.2e}", flush=True)
    assert rel < 5e-5
    assert err < 1e-5


@pytest.mark.parametrize("depth", [0, 2, 4])
@pytest.mark.parametrize("grid", ["even", "odd"])
def test_batched_anderson_matches_oracle(gpu, grid, depth):
    """test_convergence (tests/test_lippmann_schwinger.py:64-99) settings, all through the
    batched device solver: iteration bounds, and strain parity with the FP64 oracle."""
    gs = py_grid(grid)
    lmbda, mu = random_material(gs)
    lam_t = torch.from_numpy(lmbda).to(gpu)
    mu_t = torch.from_numpy(mu).to(gpu)
    s = LSSolver(gs, nrhs=2, depth=depth)
    eb = np.stack([EPS_BAR, -0.5 * EPS_BAR[::-1]]).astype(np.float32)
    out = s.solve(lam_t, mu_t, eb, rtol=1e-20, atol=1e-5, maxiter=32, want_sigma=True)
    its = out["iters"]
    for r in range(2):
        e_ref, s_ref, it_ref, _ = O.lippmann_schwinger_rfft(lmbda.astype(np.float64), mu.astype(np.float64),
                                                            eb[r].astype(np.float64), gs, rtol=1e-20, atol=1e-5,
                                                            depth=depth, maxiter=32)
        e_k, s_k, _ = oracle_fixed_updates(lmbda, mu, eb[r].astype(np.float64), gs, int(its[r]), depth=depth)
        err = rel_l2(out["epsilon"][r].cpu().numpy(), e_k)
        print(f"AA depth={depth} {grid} rhs{r}: iter gpu={its[r]} oracle={it_ref} strain relL2={err:.2e}",
              flush=True)
        assert its[r] < 8  # float32 bound of test_convergence
        assert abs(int(its[r]) - it_ref) <= 1
        assert err <= 1e-5


def test_effective_tensor_config1(gpu):
    """32^3 cross-ply (4 layers x 8 voxels, r = 2.5, vf 0.4 RSA): C_eff from one batched
    solve of the six load cases against the FP64 oracle."""
    from jaxmaterials_amd.data import crossply_lame

    shape = (32, 32, 32)
    gs = GridSpec(*shape, 32.0, 32.0, 32.0)
    lam, mu, vf, nf = crossply_lame(shape, (32.0, 32.0, 32.0), 4, 2.5, 0.4, seed=1)
    lam_np, mu_np = lam.cpu().numpy(), mu.cpu().numpy()
    C, its = effective_tensor(lam_np, mu_np, gs, rtol=1e-6, maxiter=2000)
    C_ref, its_ref = O.effective_tensor(lam_np.astype(np.float64), mu_np.astype(np.float64), gs, rtol=1e-6)
    rel = np.abs(C - C_ref).max() / np.abs(C_ref).max()
    print(f"config1 C_eff: vf={vf:.3f} fibres={nf} iters gpu={its.tolist()} oracle={its_ref} "
          f"max rel err={rel:.2e}", flush=True)
    assert rel <= 1e-5
    assert all(abs(int(a) - b) <= 1 for a, b in zip(its, its_ref))


def test_rasteriser_matches_oracle(gpu):
    from jaxmaterials_amd.data import crossply_lame
    from oracle import microstructure_oracle as MO

    for shape, nl, r, vf, seed in [((32, 32, 32), 4, 2.5, 0.4, 1), ((48, 40, 36), 3, 4.0, 0.5, 5),
                                   ((64, 64, 64), 8, 5.0, 0.6, 9)]:
        ext = tuple(float(n) for n in shape)
        lam, mu, avf, nf = crossply_lame(shape, ext, nl, r, vf, seed=seed)
        lam_r, mu_r, vf_r, fib = MO.crossply(shape, ext, nl, r, vf, seed)
        assert nf == len(fib)
        assert np.array_equal(lam.cpu().numpy(), lam_r)
        assert np.array_equal(mu.cpu().numpy(), mu_r)
        assert abs(avf - vf_r) < 1e-12


@pytest.mark.parametrize("n", [64])
def test_fp32_noise_floor_crossply(gpu, n):
    """At rtol 1e-6 the FP32 iteration must track the FP64 oracle's iteration count
    (+-1): the correction formulation keeps FP32 FFT rounding of the large base
    stress out of the residual.  64^3 cross-ply, fibre radius 1.5, 4 layers."""
    from jaxmaterials_amd.data import crossply_lame

    shape = (n, n, n)
    gs = GridSpec(*shape, float(n), float(n), float(n))
    lam, mu, vf, nf = crossply_lame(shape, (float(n),) * 3, max(2, n // 16), 6.0 * n / 256, 0.5, seed=2)
    s = LSSolver(gs, nrhs=1)
    e0 = np.zeros((1, 6), np.float32)
    e0[0, 0] = 1.0
    out = s.solve(lam, mu, e0, rtol=1e-6, maxiter=1000)
    lam_np, mu_np = lam.cpu().numpy().astype(np.float64), mu.cpu().numpy().astype(np.float64)
    e_ref, _, it_ref, _ = O.lippmann_schwinger_rfft(lam_np, mu_np, e0[0].astype(np.float64), gs, rtol=1e-6,
                                                    maxiter=1000, workers=8)
    err = rel_l2(out["epsilon"][0].cpu().numpy(), e_ref)
    print(f"crossply {n}^3: iter gpu={out['iters'][0]} oracle={it_ref} strain relL2={err:.2e}", flush=True)
    assert abs(int(out["iters"][0]) - it_ref) <= 1
    assert err <= 1e-5


# Axis lengths with register two-stage FFT kernels in the y and x passes
# (256, 128, 96, 64, 48, 32), and the same lengths on z; each grid keeps the
# other axes short so the FP64 oracle stays cheap.
LONG_AXES = [(256, 96, 8), (96, 256, 8), (128, 48, 10), (48, 128, 10), (32, 64, 12), (64, 32, 12),
             (12, 10, 256), (10, 12, 128), (16, 12, 96), (10, 14, 48)]


@pytest.mark.parametrize("shape", LONG_AXES, ids=lambda s: "x".join(map(str, s)))
def test_register_fft_lengths_match_oracle(gpu, shape):
    """Strain after a fixed number of plain and Anderson updates against the FP64
    oracle, on grids whose x and y lengths select the register FFT kernels."""
    gs = GridSpec(*shape, 1.0, 0.9, 0.3)
    lmbda, mu = random_material(gs, seed=17, lo=0.5, hi=1.5)
    lam_t = torch.from_numpy(lmbda).to(gpu)
    mu_t = torch.from_numpy(mu).to(gpu)
    for depth in (0, 3):
        s = LSSolver(gs, nrhs=1, depth=depth)
        eb = EPS_BAR[None].astype(np.float32)
        out = s.solve(lam_t, mu_t, eb, rtol=1e-20, atol=0.0, maxiter=6, check=False)
        e_k, _, _ = oracle_fixed_updates(lmbda, mu, EPS_BAR, gs, 6, depth=depth)
        err = rel_l2(out["epsilon"][0].cpu().numpy(), e_k)
        print(f"register FFT {shape} depth {depth}: strain relL2 after 6 updates={err:.2e}", flush=True)
        assert err <= 1e-5
        s.close()


# Axis lengths with a prime factor above 5 next to factors 2, 3 or 5 (the LDS
# engine's butterfly stages followed by one generic stage), and a length that
# is a product of such primes only (one direct DFT).
MIXED_AXES = [(22, 21, 26), (55, 14, 9), (39, 34, 8), (77, 6, 10)]


@pytest.mark.parametrize("shape", MIXED_AXES, ids=lambda s: "x".join(map(str, s)))
def test_mixed_radix_lengths_match_oracle(gpu, shape):
    """Strain after a fixed number of plain and Anderson updates against the FP64
    oracle on grids whose axis lengths need the generic FFT stage."""
    gs = GridSpec(*shape, 1.0, 0.9, 0.3)
    lmbda, mu = random_material(gs, seed=23, lo=0.5, hi=1.5)
    lam_t = torch.from_numpy(lmbda).to(gpu)
    mu_t = torch.from_numpy(mu).to(gpu)
    for depth in (0, 3):
        s = LSSolver(gs, nrhs=1, depth=depth)
        eb = EPS_BAR[None].astype(np.float32)
        out = s.solve(lam_t, mu_t, eb, rtol=1e-20, atol=0.0, maxiter=6, check=False)
        e_k, _, _ = oracle_fixed_updates(lmbda, mu, EPS_BAR, gs, 6, depth=depth)
        err = rel_l2(out["epsilon"][0].cpu().numpy(), e_k)
        print(f"mixed-radix FFT {shape} depth {depth}: strain relL2 after 6 updates={err:.2e}", flush=True)
        assert err <= 1e-5
        s.close()


def _register_configs():
    """(kind, n, index) of every compiled register FFT configuration, read from
    the LS_REG_Y_LIST / LS_REG_H_LIST tables of the solver source (the index
    counts the configurations of one length in table order)."""
    import os
    import re

    src = open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "jaxmaterials_amd",
                            "csrc", "ls_solver.hip")).read()
    out = []
    for kind, macro in (("y", "LS_REG_Y_LIST"), ("x", "LS_REG_H_LIST")):
        body = re.search(r"#define " + macro + r"\(F\)((?:[^\n]*\\\n)*[^\n]*)", src).group(1)
        seen = {}
        for m in re.finditer(r"F\((\d+),", body):
            n = int(m.group(1))
            out.append((kind, n, seen.get(n, 0)))
            seen[n] = seen.get(n, 0) + 1
    return out


REG_CONFIGS = _register_configs()


def test_register_config_table_parsed():
    assert len(REG_CONFIGS) >= 30
    assert ("y", 160, 0) in REG_CONFIGS and ("x", 256, 0) in REG_CONFIGS


@pytest.mark.parametrize("cfg", REG_CONFIGS, ids=lambda c: f"{c[0]}{c[1]}_{c[2]}")
def test_every_register_config_matches_oracle(gpu, cfg, monkeypatch):
    """Each compiled register configuration (not only a length's default),
    selected by LS_REG_Y_INDEX / LS_REG_X_INDEX, against the FP64 oracle after
    six plain and six Anderson(3) updates."""
    kind, n, idx = cfg
    shape = (10, n, 8) if kind == "y" else (n, 10, 8)
    monkeypatch.setenv("LS_REG_Y_INDEX" if kind == "y" else "LS_REG_X_INDEX", str(idx))
    gs = GridSpec(*shape, 1.0, 0.9, 0.3)
    lmbda, mu = random_material(gs, seed=29, lo=0.5, hi=1.5)
    lam_t = torch.from_numpy(lmbda).to(gpu)
    mu_t = torch.from_numpy(mu).to(gpu)
    for depth in (0, 3):
        s = LSSolver(gs, nrhs=1, depth=depth)
        out = s.solve(lam_t, mu_t, EPS_BAR[None].astype(np.float32), rtol=1e-20, atol=0.0, maxiter=6, check=False)
        e_k, _, _ = oracle_fixed_updates(lmbda, mu, EPS_BAR, gs, 6, depth=depth)
        err = rel_l2(out["epsilon"][0].cpu().numpy(), e_k)
        print(f"register {kind} n={n} config {idx} depth {depth}: strain relL2 {err:.2e}", flush=True)
        assert err <= 1e-5
        s.close()


def test_effective_tensor_256_against_fp64_oracle(gpu):
    """The benchmark workload: 256^3 cross-ply (16 layers of 16 voxels, fibre
    radius 6, RSA vf 0.5, seed 2), six load cases batched, Anderson depth 6 (the
    benchmark's), depth 5 and plain iteration to rtol 1e-6.  C_eff against the FP64 oracle's (Anderson
    depth 5 to rtol 1e-9, tests/golden/crossply256_ceff.json): within 1e-5 of
    the largest entry, and within 1e-5 relative on every entry above 1e-3 of
    the largest."""
    import json
    import os

    from jaxmaterials_amd.data import crossply_lame

    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "crossply256_ceff.json")
    with open(path) as f:
        gold = json.load(f)
    cfg = gold["config"]
    shape = tuple(cfg["shape"])
    gs = GridSpec(*shape, *cfg["extents"])
    lam, mu, vf, nf = crossply_lame(shape, tuple(cfg["extents"]), cfg["nlayers"], cfg["radius"], cfg["vf"],
                                    seed=cfg["seed"])
    assert nf == gold["n_fibres"] and abs(vf - gold["achieved_vf"]) < 1e-12
    G = np.array(gold["ceff"])
    sig = np.abs(G) > 1e-3 * np.abs(G).max()
    for depth in (6, 5, 0):
        s = LSSolver(gs, nrhs=6, depth=depth)
        out = s.solve(lam, mu, unit_load_cases(), rtol=1e-6, maxiter=3000)
        C = out["sigma_mean"]
        norm_err = np.abs(C - G).max() / np.abs(G).max()
        entry_err = (np.abs(C - G)[sig] / np.abs(G)[sig]).max()
        print(f"256^3 C_eff depth {depth}: iters {out['iters'].tolist()} max err / max entry {norm_err:.2e}, "
              f"max entry relative err {entry_err:.2e}", flush=True)
        assert norm_err <= 1e-5
        assert entry_err <= 1e-5
        s.close()
        del s
        torch.cuda.empty_cache()


@pytest.mark.parametrize("shape", [(24, 20, 4), (16, 18, 6), (20, 12, 5)], ids=lambda s: "x".join(map(str, s)))
def test_anderson_gram_on_self_conjugate_planes(gpu, shape):
    """Anderson trajectories on grids where the self-conjugate planes k_z = 0
    and nz/2 hold most of the modes: the Gram row (spectral form plus the
    plane correction recorded by the inverse y pass) must be the real-space
    one, so the strain after a fixed number of Anderson updates matches the
    FP64 oracle, whose Gram row is formed in real space
    (lippmann_schwinger.py:217-224)."""
    gs = GridSpec(*shape, 1.0, 0.8, 0.5)
    lmbda, mu = random_material(gs, seed=23, lo=0.3, hi=1.7)
    lam_t = torch.from_numpy(lmbda).to(gpu)
    mu_t = torch.from_numpy(mu).to(gpu)
    s = LSSolver(gs, nrhs=1, depth=4)
    out = s.solve(lam_t, mu_t, EPS_BAR[None].astype(np.float32), rtol=1e-30, atol=0.0, maxiter=10, check=False)
    e_k, _, _ = oracle_fixed_updates(lmbda, mu, EPS_BAR, gs, 10, depth=4)
    err = rel_l2(out["epsilon"][0].cpu().numpy(), e_k)
    print(f"Anderson planes {shape}: strain relL2 after 10 updates={err:.2e}", flush=True)
    assert err <= 1e-5
    s.close()
