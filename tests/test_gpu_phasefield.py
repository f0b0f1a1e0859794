"""Phase-field fracture on the GPU against the FP64 oracle
(oracle/phasefield_oracle.py): the damaged stress and the strain energy
pointwise, the damaged Lippmann-Schwinger solve (fixed update counts and
converged, plain and Anderson, 3-D grids and the nz = 1 grid of the
reference's script), the FP64 phase-field fixed point, and a few staggered
load steps."""

import numpy as np
import pytest
import torch

from jaxmaterials_amd.common import GridSpec
from oracle import phasefield_oracle as PF

pytestmark = pytest.mark.gpu


def _fields(gs, seed, dmax=0.9):
    rng = np.random.default_rng(seed)
    shape = (gs.nx, gs.ny, gs.nz)
    lam = rng.uniform(0.8, 1.6, shape).astype(np.float32)
    mu = rng.uniform(0.6, 1.2, shape).astype(np.float32)
    d = rng.uniform(0.0, dmax, shape).astype(np.float32)
    return lam, mu, d


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_damaged_stress_and_energy_pointwise(gpu):
    from jaxmaterials_amd.solver.phasefield import compute_sigma_damaged, compute_strain_energy

    rng = np.random.default_rng(1)
    shape = (9, 8, 7)
    lam = rng.uniform(0.5, 2.0, shape).astype(np.float32)
    mu = rng.uniform(0.5, 2.0, shape).astype(np.float32)
    eps = rng.normal(0, 1e-2, (6,) + shape).astype(np.float32)
    eps[:, 0, 0, :] = 0.0  # zero strain
    eps[3:, 1, :, :] = 0.0  # axis-aligned states (diagonal matrices)
    d = rng.uniform(0, 1, shape).astype(np.float32)
    t = lambda a: torch.from_numpy(a).to(gpu)  # noqa: E731
    s = compute_sigma_damaged(t(lam), t(mu), t(eps), t(d), 1e-6).cpu().numpy()
    want = PF.sigma_damaged(lam, mu, eps, d, 1e-6)
    assert np.abs(s - want).max() <= 2e-7 * np.abs(want).max()
    s0 = compute_sigma_damaged(t(lam), t(mu), t(eps)).cpu().numpy()
    assert np.abs(s0 - PF.sigma_damaged(lam, mu, eps)).max() <= 2e-7 * np.abs(want).max()
    psi = compute_strain_energy(t(lam), t(mu), t(eps)).cpu().numpy()
    wpsi = PF.strain_energy_plus(lam, mu, eps)
    assert np.abs(psi - wpsi).max() <= 1e-12 * np.abs(wpsi).max() + 1e-9 * np.abs(wpsi).max()
    H = torch.full(shape, 0.5 * float(wpsi.max()), dtype=torch.float64, device=gpu)
    compute_strain_energy(t(lam), t(mu), t(eps), history=H)
    assert np.allclose(H.cpu().numpy(), np.maximum(0.5 * wpsi.max(), wpsi), rtol=1e-9, atol=0)


GRIDS = [((12, 10, 8), (1.0, 0.9, 0.7)), ((20, 18, 1), (1.0, 1.0, 0.015))]


@pytest.mark.parametrize("shape,extents", GRIDS, ids=["12x10x8", "20x18x1"])
@pytest.mark.parametrize("depth", [0, 3])
def test_damaged_solve_matches_oracle(gpu, shape, extents, depth):
    from jaxmaterials_amd.solver.phasefield import elasticity_solve

    gs = GridSpec(*shape, *extents)
    lam, mu, d = _fields(gs, seed=7)
    eb = np.array([0.002, 0.01, -0.001, 0.002, 0.0, 0.001])
    t = lambda a: torch.from_numpy(a).to(gpu)  # noqa: E731
    # the trajectory: a fixed number of updates
    e_g, s_g, it_g = elasticity_solve(t(lam), t(mu), eb, t(d), 1e-6, gs, rtol=1e-20, atol=0.0, depth=depth, maxiter=8)
    e_o, s_o, it_o = PF.elasticity_solve(lam, mu, eb, d, 1e-6, gs, rtol=1e-20, atol=0.0, depth=depth, maxiter=8)
    err = _rel(e_g.cpu().numpy(), e_o)
    print(f"damaged LS {shape} depth {depth}: 8 updates strain relL2 {err:.2e}", flush=True)
    assert it_g == it_o == 8
    assert err <= 1e-5
    # converged: iterations within one, strain and stress to 1e-5
    e_g, s_g, it_g = elasticity_solve(t(lam), t(mu), eb, t(d), 1e-6, gs, rtol=1e-6, depth=depth, maxiter=1000)
    e_o, s_o, it_o = PF.elasticity_solve(lam, mu, eb, d, 1e-6, gs, rtol=1e-6, depth=depth, maxiter=1000)
    err_e, err_s = _rel(e_g.cpu().numpy(), e_o), _rel(s_g.cpu().numpy(), s_o)
    print(f"damaged LS {shape} depth {depth}: iterations gpu {it_g} oracle {it_o}, strain {err_e:.2e} "
          f"stress {err_s:.2e}", flush=True)
    assert abs(it_g - it_o) <= 1
    assert err_e <= 1e-5 and err_s <= 1e-5


def test_undamaged_law_is_the_linear_solve(gpu):
    from jaxmaterials_amd.solver.device import LSSolver
    from jaxmaterials_amd.solver.phasefield import elasticity_solve

    gs = GridSpec(12, 10, 8, 1.0, 0.9, 0.7)
    lam, mu, _ = _fields(gs, seed=3)
    eb = np.array([0.01, 0.0, 0.0, 0.0, 0.003, 0.0])
    t = lambda a: torch.from_numpy(a).to(gpu)  # noqa: E731
    e1, _, it1 = elasticity_solve(t(lam), t(mu), eb, torch.zeros_like(t(lam)), 0.0, gs, rtol=1e-6, depth=2,
                                  maxiter=500)
    s = LSSolver(gs, nrhs=1, depth=2)
    out = s.solve(t(lam), t(mu), eb[None].astype(np.float32), rtol=1e-6, atol=1e-20, maxiter=500)
    assert it1 == int(out["iters"][0])
    assert _rel(e1.cpu().numpy(), out["epsilon"][0].cpu().numpy()) <= 1e-6


@pytest.mark.parametrize("shape,extents", [((12, 10, 8), (1.0, 0.9, 0.7)), ((55, 50, 1), (1.0, 1.0, 0.015))],
                         ids=["12x10x8", "55x50x1"])
def test_phase_field_solve_matches_oracle(gpu, shape, extents):
    from jaxmaterials_amd.solver.phasefield import phase_field_solve

    gs = GridSpec(*shape, *extents)
    rng = np.random.default_rng(9)
    H = rng.uniform(0.0, 0.05, shape)
    d_old = rng.uniform(0.0, 0.3, shape)
    gc, lc = 2.7e-3, 0.015
    d_g, it_g = phase_field_solve(torch.from_numpy(H).to(gpu), torch.from_numpy(d_old).to(gpu), gc, lc, gs,
                                  tolerance=1e-9, maxiter=10000)
    d_o, it_o = PF.phase_field_solve(H, d_old, gc, lc, gs, tolerance=1e-9, maxiter=10000)
    err = np.abs(d_g.cpu().numpy() - d_o).max() / np.abs(d_o).max()
    print(f"phase field {shape}: iterations gpu {it_g} oracle {it_o}, max rel err {err:.2e}", flush=True)
    assert abs(it_g - it_o) <= 1
    # FP64 FFTs on both sides, the symbol from the FP64 extents
    assert err <= 1e-12


def test_staggered_steps_match_oracle(gpu):
    from jaxmaterials_amd.solver.phasefield import staggered_steps

    gs = GridSpec(20, 18, 1, 1.0, 1.0, 0.015)
    shape = (20, 18, 1)
    lam = np.full(shape, 121.15, dtype=np.float32)
    mu = np.full(shape, 80.77, dtype=np.float32)
    # a notch: a soft strip from the left edge, as the reference script's crack
    lam[:8, 9, :] = 1e-3
    mu[:8, 9, :] = 1e-3
    ebars = [np.array([0.0, e, 0.0, 0.0, 0.0, 0.0]) for e in np.linspace(0.0, 0.01, 5)]
    gc, lc, k = 2.7e-3, 0.05, 1e-6
    out = staggered_steps(torch.from_numpy(lam).to(gpu), torch.from_numpy(mu).to(gpu), ebars, gs, gc, lc, k,
                          depth=3, rtol=1e-6, maxiter=1000, pf_tolerance=1e-9, pf_maxiter=10000)
    # the same loop on the oracle
    d = np.zeros(shape)
    H = np.zeros(shape)
    for i, eb in enumerate(ebars):
        d, _ = PF.phase_field_solve(H, d, gc, lc, gs, tolerance=1e-9, maxiter=10000)
        eps, sig, it = PF.elasticity_solve(lam, mu, eb, d.astype(np.float32), k, gs, rtol=1e-6, depth=3,
                                           maxiter=1000)
        H = np.maximum(H, PF.strain_energy_plus(lam, mu, eps))
        st = out["steps"][i]
        sm = sig.mean(axis=(1, 2, 3))
        print(f"step {i}: iterations gpu {st['elasticity_iterations']} oracle {it}, sigma_yy gpu "
              f"{st['sigma_mean'][1]:.6e} oracle {sm[1]:.6e}", flush=True)
        # the notch makes the contrast 1e5: with Anderson mixing the FP32
        # trajectory reaches the stop test within a couple of updates of the
        # FP64 one (the tensors below agree to 1e-5)
        assert abs(st["elasticity_iterations"] - it) <= max(2, 0.1 * it)
        assert abs(st["sigma_mean"][1] - sm[1]) <= 1e-5 * max(abs(sm[1]), 1e-12)
    assert np.abs(out["damage"].cpu().numpy() - d).max() <= 1e-5 * max(np.abs(d).max(), 1e-12)
    assert np.abs(out["history"].cpu().numpy() - H).max() <= 1e-5 * np.abs(H).max()
    assert np.abs(d).max() > 0.0


@pytest.mark.parametrize("maxiter", [0, 1, 5, 8, 13])
def test_phase_field_iteration_cap(gpu, maxiter):
    """The fixed point replays chunks of iterations with the stop test on the
    device: a cap that is not a multiple of the chunk, a cap of one, and no
    iteration at all must give the oracle's count and field."""
    from jaxmaterials_amd.solver.phasefield import phase_field_solve

    gs = GridSpec(12, 10, 8, 1.0, 0.9, 0.7)
    rng = np.random.default_rng(4)
    H = rng.uniform(0.0, 0.05, (12, 10, 8))
    d_old = rng.uniform(0.0, 0.3, (12, 10, 8))
    d_g, it_g = phase_field_solve(torch.from_numpy(H).to(gpu), torch.from_numpy(d_old).to(gpu), 2.7e-3, 0.015, gs,
                                  tolerance=1e-30, maxiter=maxiter)
    d_o, it_o = PF.phase_field_solve(H, d_old, 2.7e-3, 0.015, gs, tolerance=1e-30, maxiter=maxiter)
    assert it_g == it_o == maxiter
    assert np.abs(d_g.cpu().numpy() - d_o).max() <= 1e-12 * max(np.abs(d_o).max(), 1.0)
