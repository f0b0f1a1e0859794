"""Batched solves of uneven microstructure families, graph replay and the tile
autotuner (through the C ABI), against single solves and the FP64 oracle."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _family_samples(grids, per_grid):
    from jaxmaterials_amd.data.batch import sample_params

    out, i = [], 0
    counts = {g: 0 for g in grids}
    while any(c < per_grid for c in counts.values()):
        p = sample_params(i)
        if p["grid"] in counts and counts[p["grid"]] < per_grid:
            out.append(p)
            counts[p["grid"]] += 1
        i += 1
    return out


def _rel_max(a, b):
    sig = np.abs(b) > 1e-3 * np.abs(b).max()
    return float((np.abs(a - b)[sig] / np.abs(b)[sig]).max())


def test_graph_replay_matches_launch_by_launch(gpu):
    """The same solve with and without graph replay, alternating, gives
    identical iterations and bit-identical C_eff on every replay."""
    from jaxmaterials_amd.common import GridSpec
    from jaxmaterials_amd.data import crossply_lame
    from jaxmaterials_amd.solver.device import LSSolver, unit_load_cases

    n = 32
    lam, mu, _, _ = crossply_lame((n,) * 3, (float(n),) * 3, 4, 2.5, 0.4, seed=11)
    solver = LSSolver(GridSpec(n, n, n, float(n), float(n), float(n)), nrhs=6, depth=4)
    runs = [solver.solve(lam, mu, unit_load_cases(), rtol=1e-6, maxiter=2000, graph=g)
            for g in (False, True, False, True, True)]
    for r in runs[1:]:
        assert np.array_equal(r["iters"], runs[0]["iters"])
        assert np.array_equal(r["sigma_mean"], runs[0]["sigma_mean"])
    solver.close()


def test_autotuned_tiles_give_the_same_result(gpu):
    from jaxmaterials_amd.common import GridSpec
    from jaxmaterials_amd.data import crossply_lame
    from jaxmaterials_amd.solver.device import LSSolver, unit_load_cases

    n = 48
    lam, mu, _, _ = crossply_lame((n,) * 3, (float(n),) * 3, 3, 3.0, 0.4, seed=12)
    gs = GridSpec(n, n, n, float(n), float(n), float(n))
    s0 = LSSolver(gs, nrhs=6, depth=4)
    ref = s0.solve(lam, mu, unit_load_cases(), rtol=1e-6, maxiter=2000)
    s1 = LSSolver(gs, nrhs=6, depth=4)
    report = s1.autotune(lam, mu)
    print("autotune 48^3:", report)
    out = s1.solve(lam, mu, unit_load_cases(), rtol=1e-6, maxiter=2000)
    assert np.all(np.abs(out["iters"] - ref["iters"]) <= 1)
    assert _rel_max(out["sigma_mean"], ref["sigma_mean"]) < 1e-5
    s0.close()
    s1.close()


def test_uneven_batch_matches_single_solves_and_oracle(gpu):
    """Samples of the uneven family on 32^3 and 48^3 (radix 3), batched with
    graph replay and autotuned tiles, against one solve per microstructure and
    (one 48^3 sample) against the FP64 oracle."""
    from jaxmaterials_amd.common import GridSpec
    from jaxmaterials_amd.data.batch import lame_fields, solve_batch
    from jaxmaterials_amd.solver.device import LSSolver, unit_load_cases
    from oracle import ls_oracle as O

    params = _family_samples((32, 48), 3)
    res = solve_batch(params, depth=4, graph=True, autotune=True, log=print)
    checked = False
    for j, p in enumerate(params):
        n = p["grid"]
        lam, mu, _, _ = lame_fields(p, device=gpu)
        s = LSSolver(GridSpec(n, n, n, float(n), float(n), float(n)), nrhs=6, depth=4)
        one = s.solve(lam, mu, unit_load_cases(), rtol=1e-6, maxiter=3000)
        s.close()
        d = _rel_max(res["ceff"][j], one["sigma_mean"])
        print(f"sample {p['index']} grid {n} ratio {p['mu_ratio']:.1f}: iters {res['iters'][j].tolist()} "
              f"single {one['iters'].tolist()} C_eff rel diff {d:.2e}")
        assert np.all(np.abs(res["iters"][j] - one["iters"]) <= 1)
        assert d < 1e-5
        if n == 48 and not checked:
            checked = True
            gs = O.GridSpec(n, n, n, float(n), float(n), float(n))
            C, _ = O.effective_tensor(lam.cpu().numpy().astype(np.float64), mu.cpu().numpy().astype(np.float64),
                                      gs, rtol=1e-8, depth=4, maxiter=3000)
            e = _rel_max(res["ceff"][j], np.asarray(C))
            print(f"48^3 sample vs FP64 oracle: C_eff max rel err {e:.2e}")
            assert e < 1e-5


def test_graph_solve_after_autotune_recaptures(gpu):
    """A graph solve, the autotuner (which changes the tiles and the base
    divergence layout) and a second graph solve on the same buffers: the second
    must equal a launch-by-launch solve with the new tiles, not replay the
    chunk captured with the old ones."""
    import torch

    from jaxmaterials_amd.common import GridSpec
    from jaxmaterials_amd.data import crossply_lame
    from jaxmaterials_amd.solver.device import LSSolver, unit_load_cases

    n = 48
    lam, mu, _, _ = crossply_lame((n,) * 3, (float(n),) * 3, 3, 3.0, 0.4, seed=21)
    s = LSSolver(GridSpec(n, n, n, float(n), float(n), float(n)), nrhs=6, depth=4)
    eps = torch.empty((6, 6, n, n, n), dtype=torch.float32, device=gpu)
    first = s.solve(lam, mu, unit_load_cases(), epsilon=eps, rtol=1e-6, maxiter=2000, graph=True)
    report = s.autotune(lam, mu)
    print("autotune 48^3:", report)
    again = s.solve(lam, mu, unit_load_cases(), epsilon=eps, rtol=1e-6, maxiter=2000, graph=True)
    plain = s.solve(lam, mu, unit_load_cases(), rtol=1e-6, maxiter=2000, graph=False)
    assert np.array_equal(again["iters"], plain["iters"])
    assert np.array_equal(again["sigma_mean"], plain["sigma_mean"])
    assert np.all(np.abs(again["iters"] - first["iters"]) <= 1)
    assert _rel_max(again["sigma_mean"], first["sigma_mean"]) < 1e-5
    s.close()


def test_uneven_batch_large_grids_match_single_solves(gpu):
    """One member each of the family's 64^3, 96^3 (radix 3) and 128^3 grids at
    the configuration's Anderson depth 4, batched (graph replay, autotuned
    tiles), against one launch-by-launch solve per microstructure."""
    from jaxmaterials_amd.common import GridSpec
    from jaxmaterials_amd.data.batch import lame_fields, solve_batch
    from jaxmaterials_amd.solver.device import LSSolver, unit_load_cases

    params = _family_samples((64, 96, 128), 1)
    res = solve_batch(params, depth=4, graph=True, autotune=True, log=print)
    for j, p in enumerate(params):
        n = p["grid"]
        lam, mu, _, _ = lame_fields(p, device=gpu)
        s = LSSolver(GridSpec(n, n, n, float(n), float(n), float(n)), nrhs=6, depth=4)
        one = s.solve(lam, mu, unit_load_cases(), rtol=1e-6, maxiter=3000)
        s.close()
        d = _rel_max(res["ceff"][j], one["sigma_mean"])
        print(f"sample {p['index']} grid {n} ratio {p['mu_ratio']:.1f}: iters {res['iters'][j].tolist()} "
              f"single {one['iters'].tolist()} C_eff rel diff {d:.2e}", flush=True)
        assert np.all(np.abs(res["iters"][j] - one["iters"]) <= 1)
        assert d < 1e-5
