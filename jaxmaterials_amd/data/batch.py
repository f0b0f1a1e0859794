"""Uneven batches of seeded cross-ply microstructures and their effective
tensors, solved many at a time on one GPU.

The family: sample i draws, from a splitmix64 stream seeded by i, a grid
uniform over {32, 48, 64, 96, 128}^3 (48 and 96 exercise radix-3 FFT stages),
2-8 layers, fibre radius uniform in [1.5, 5] voxels, volume fraction uniform
in [0.3, 0.6], fibre/matrix shear-modulus ratio log-uniform in [5, 100];
matrix mu = 1.111, nu = 0.35; fibre nu = 0.2; six unit load cases each.

Samples are grouped by grid and solved in batches (all load cases of up to
`max_microstructures` microstructures per solver call, Anderson depth 4 as the
family's configuration asks, or conjugate gradients), with the tiles of each
grid picked by the solver's autotuner, the iterations between host
synchronisations replayed as captured HIP graphs, and the five grid groups
running concurrently on their own streams.
This replaces the reference's one-call-per-load-case loop
(src/jaxmaterials/data/data.py:148-186).

    python -m jaxmaterials_amd.data.batch --count 384
"""

import argparse
import json
import math
import sys
import time

import numpy as np
import torch

from ..common import GridSpec
from ..solver.device import LSSolver
from .microstructure import crossply_lame

__all__ = ["GRIDS", "sample_params", "solve_batch"]

GRIDS = (32, 48, 64, 96, 128)
_MASK = (1 << 64) - 1


def _splitmix64(state):
    """One step of splitmix64 (Steele, Lea & Flood 2014): (new state, output)."""
    state = (state + 0x9E3779B97F4A7C15) & _MASK
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
    return state, z ^ (z >> 31)


def sample_params(i, family_seed=0x3C0F16):
    """Parameters of sample i of the family (deterministic, platform independent)."""
    st = (family_seed * 0x100000001B3 + i) & _MASK
    u = []
    for _ in range(6):
        st, z = _splitmix64(st)
        u.append((z >> 11) * (1.0 / (1 << 53)))
    mu_m, nu_m, nu_f = 1.111, 0.35, 0.2
    ratio = math.exp(math.log(5.0) + (math.log(100.0) - math.log(5.0)) * u[4])
    mu_f = ratio * mu_m
    return {
        "index": i,
        "grid": GRIDS[min(int(u[0] * len(GRIDS)), len(GRIDS) - 1)],
        "layers": 2 + min(int(u[1] * 7), 6),
        "radius": 1.5 + 3.5 * u[2],
        "vf": 0.3 + 0.3 * u[3],
        "mu_ratio": ratio,
        "seed": int(z),
        "lam_matrix": 2.0 * mu_m * nu_m / (1.0 - 2.0 * nu_m),
        "mu_matrix": mu_m,
        "lam_fibre": 2.0 * mu_f * nu_f / (1.0 - 2.0 * nu_f),
        "mu_fibre": mu_f,
    }


def lame_fields(p, device="cuda"):
    n = p["grid"]
    lam, mu, vf, nf = crossply_lame((n, n, n), (float(n),) * 3, p["layers"], p["radius"], p["vf"], p["seed"],
                                    lam_matrix=p["lam_matrix"], mu_matrix=p["mu_matrix"],
                                    lam_fibre=p["lam_fibre"], mu_fibre=p["mu_fibre"], device=device)
    return lam, mu, vf, nf


def _batch_size(n, depth, budget_bytes, cg=False):
    """Microstructures per solver call within a device-memory budget
    (about 9 (m + 1) + 21 fields of N floats per load case, m = depth: base,
    correction, half spectrum, base divergence, and the Anderson rings of
    g (6 N per slot) and of the spectral residual form (3 N per slot);
    conjugate gradients: 6 more for the direction instead of the rings;
    a margin of 3 fields is added)."""
    per_rhs = ((30 if cg else 9 * (depth + 1) + 24)) * n**3 * 4
    return max(1, int(budget_bytes // (6 * per_rhs)))


def _solve_grid(n, idx, params, depth, rtol, maxiter, graph, autotune, budget_bytes, max_microstructures, dev,
                cg, out_c, out_it, log, sync):
    """All microstructures of one grid size, in calls of up to k microstructures
    on one solver (its own HIP stream).  sync: (lock, barrier, start) shared by
    the grid threads, so that the autotuner runs alone and every thread starts
    solving after all have tuned (start[0] = that time).  Returns the grid's
    timing record."""
    lock, barrier, start = sync
    k = min(len(idx), _batch_size(n, depth, budget_bytes, cg), max_microstructures)
    gs = GridSpec(n, n, n, float(n), float(n), float(n))
    solver = LSSolver(gs, nrhs=6 * k, nmat=k, depth=depth)
    lam = torch.empty((k, n, n, n), dtype=torch.float32, device=dev)
    mu = torch.empty_like(lam)
    eps = torch.empty((6 * k, 6, n, n, n), dtype=torch.float32, device=dev)
    eb = np.tile(np.eye(6, dtype=np.float32), (k, 1))
    mats = np.repeat(np.arange(k, dtype=np.int32), 6)
    report = None
    t_solve = t_gen = 0.0
    ncalls = 0
    t_all = time.perf_counter()
    for c0 in range(0, len(idx), k):
        chunk = idx[c0:c0 + k]
        tg = time.perf_counter()
        for q in range(k):
            l, m, _, _ = lame_fields(params[chunk[min(q, len(chunk) - 1)]], device=dev)
            lam[q].copy_(l)
            mu[q].copy_(m)
        if c0 == 0:
            if autotune:
                with lock:
                    report = solver.autotune(lam, mu, epsilon=eps)
            torch.cuda.current_stream(dev).synchronize()
            if barrier.wait() == 0:
                start[0] = time.perf_counter()
            t_all = time.perf_counter()
        torch.cuda.current_stream(dev).synchronize()
        t0 = time.perf_counter()
        t_gen += t0 - tg
        res = solver.solve(lam, mu, eb, rhs_mat=mats, epsilon=eps, rtol=rtol, maxiter=maxiter, graph=graph, cg=cg)
        t_solve += time.perf_counter() - t0
        ncalls += 1
        for q, j in enumerate(chunk):
            out_c[j] = res["sigma_mean"][6 * q:6 * q + 6]
            out_it[j] = res["iters"][6 * q:6 * q + 6]
    rec = {"microstructures": len(idx), "per_call": k, "calls": ncalls, "solve_s": t_solve,
           "generate_and_tune_s": t_gen, "wall_s": time.perf_counter() - t_all,
           "iterations_max_mean": float(np.mean(out_it[idx].max(axis=1))), "autotune": report}
    if log:
        log(f"grid {n}^3: {len(idx)} microstructures in {ncalls} calls of {k}, solve {t_solve:.2f} s, "
            f"mean max iterations {rec['iterations_max_mean']:.1f}; {report}")
    solver.close()
    return rec


def solve_batch(params, depth=4, rtol=1e-6, maxiter=3000, graph=True, autotune=True, budget_bytes=40 << 30,
                max_microstructures=128, device="cuda", log=None, cg=False, concurrent=True):
    """Effective tensors of the given samples.  Returns a dict with
    ceff [n][6][6] (row k = <sigma> under eps_bar = e_k), iters [n][6],
    per-grid timings and the wall time of the solver part.

    Samples are grouped by grid size; each group has its own solver, and so
    its own HIP stream, and with ``concurrent`` the groups run at the same
    time from one host thread each (the library's calls release the GIL), so
    the latency-bound small grids overlap the bandwidth-bound large ones.
    depth = Anderson depth (the family's configuration uses 4); cg = conjugate
    gradients instead."""
    import threading

    dev = torch.device(device)
    if cg:
        depth = 0  # conjugate gradients keep no Anderson history
    out_c = np.zeros((len(params), 6, 6))
    out_it = np.zeros((len(params), 6), dtype=np.int32)
    by_grid = {}
    for j, p in enumerate(params):
        by_grid.setdefault(p["grid"], []).append(j)
    grids = sorted(by_grid, reverse=True)  # largest first: they set the critical path
    timing = {}
    errors = []

    def work(n, sync):
        try:
            torch.cuda.set_device(dev.index if dev.index is not None else torch.cuda.current_device())
            timing[n] = _solve_grid(n, by_grid[n], params, depth, rtol, maxiter, graph, autotune, budget_bytes,
                                    max_microstructures, dev, cg, out_c, out_it, log, sync)
        except Exception as exc:  # reported after the join
            errors.append((n, exc))
            sync[1].abort()

    torch.cuda.synchronize(dev)
    start = [time.perf_counter()]
    if concurrent:
        sync = (threading.Lock(), threading.Barrier(len(grids)), start)
        threads = [threading.Thread(target=work, args=(n, sync)) for n in grids]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
    else:
        for n in grids:
            work(n, (threading.Lock(), threading.Barrier(1), [0.0]))
    torch.cuda.synchronize(dev)
    # concurrent: from the moment every grid has tuned; serial: the sum of the solves
    wall = time.perf_counter() - start[0] if concurrent else sum(t["solve_s"] for t in timing.values())
    if errors:
        raise RuntimeError(f"batch solve failed on grid {errors[0][0]}: {errors[0][1]}") from errors[0][1]
    torch.cuda.empty_cache()
    return {"ceff": out_c, "iters": out_it, "timing": timing, "wall_s": wall,
            "solve_s": sum(t["solve_s"] for t in timing.values())}


def main(argv=None):
    ap = argparse.ArgumentParser(description="C_eff of an uneven batch of seeded cross-ply microstructures")
    ap.add_argument("--count", type=int, default=384)
    ap.add_argument("--depth", type=int, default=4)
    ap.add_argument("--cg", action="store_true", help="conjugate gradients instead of Anderson")
    ap.add_argument("--serial", action="store_true", help="one grid group after the other")
    ap.add_argument("--rtol", type=float, default=1e-6)
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    args = ap.parse_args(argv)
    log = lambda s: print(s, flush=True)
    params = [sample_params(i) for i in range(args.count)]
    t0 = time.perf_counter()
    res = solve_batch(params, depth=args.depth, rtol=args.rtol, graph=not args.no_graph,
                      autotune=not args.no_autotune, log=log, cg=args.cg, concurrent=not args.serial)
    wall = time.perf_counter() - t0
    line = {"microstructures": args.count, "solve_s": res["solve_s"], "wall_s": wall,
            "ceff_solves_per_s": args.count / res["wall_s"], "depth": args.depth, "cg": args.cg,
            "concurrent": not args.serial, "graph": not args.no_graph,
            "autotune": not args.no_autotune,
            "grids": {str(k): {kk: vv for kk, vv in v.items() if kk != "autotune"} for k, v in res["timing"].items()}}
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
