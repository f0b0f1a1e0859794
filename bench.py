#!/usr/bin/env python
"""C_eff solves/sec to a 1e-6 relative residual at 256^3: plain LS, batched
Anderson, and Anderson from the learned warm start.

One step = one effective-tensor solve: the six unit Voigt load cases of a
256^3 cross-ply microstructure (16 layers of 16 voxels, fibre radius 6 voxels,
RSA to volume fraction 0.5, matrix (lambda, mu) = (2.593, 1.111), fibre
(19.44, 29.17)) batched in one solver call, iterated to R <= 1e-6 R(eps_bar),
then the FP64 volume averages that form the 6x6 C_eff.  Inputs are generated
on the GPU and resident in HBM before the timed region.  The warm-started
method ("aa_ws") times the network inference and the solve together.

Prints one JSON line.  Usage: python bench.py [--steps K] [--warmup W]
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "C_eff solves/sec to 1e-6 residual at 256³ (plain vs batched AA + learned start)"


def model_bytes(kind, N, M, iters, depth, nrebase=0, cg=False):
    """Algorithmic HBM bytes of all launches of one kernel kind over a solve.

    One HBM round trip of the z-half spectrum (48 M bytes per load case) per
    pass and direction; per load case and iteration j (f_j = min(j, depth+1)
    filled Anderson slots when the z pass of iteration j runs):
      z plain    : spectrum 96 M + strain 48 N (+ Lame 8 N shared by the batch)
      z Anderson : spectrum 96 M + newest g written 24 N + older g read
                   (f_j - 1) 24 N + once the ring is full (j > depth + 1) the
                   oldest g read 24 N to rebuild the current correction (+ Lame
                   8 N shared; the first iteration after a rebase skips that
                   read, not counted)
      y (inverse + forward) : 2 x 96 M
      x : 120 M per evaluation (setup, every check, every re-evaluation): the
          spectrum read and written and the base divergence d_base read; with
          Anderson also the residual's three-complex-per-mode form y written 24 M
      aa_gram : Anderson Gram row, the f newest y read, f * 24 M,
          f = min(e + 1, depth + 1) at the e-th evaluation
    Conjugate gradients (cg):
      z_cg : spectrum 96 M + correction read and written 48 N + direction p
             read and written 48 N (+ Lame 8 N shared)
      x    : 144 M per evaluation: the spectrum read and written and the
             running divergence read and written
    """
    its = [int(i) for i in iters]
    m1 = depth + 1
    nb = max(its) if its else 0
    if kind == "z":
        return sum(k * (96 * M + 48 * N) for k in its) + nb * 8 * N if depth == 0 and not cg else 0
    if kind == "z_cg":
        return sum(k * (96 * M + 96 * N) for k in its) + nb * 8 * N if cg else 0
    if kind == "z_aa":
        return sum(96 * M + 24 * N + 24 * N * (min(j, m1) - 1) + (24 * N if j > m1 else 0)
                   for k in its for j in range(1, k + 1)) + nb * 8 * N
    if kind == "y":
        return sum(2 * k for k in its) * 96 * M
    if kind == "x":
        if cg:
            return (sum(k + 1 for k in its) + nrebase * len(its)) * 144 * M
        # spectrum in and out (96 M) and the base divergence d_base read (24 M)
        b = (sum(k + 1 for k in its) + nrebase * len(its)) * 120 * M
        if depth > 0:
            b += sum(k + 1 for k in its) * 24 * M
        return b
    if kind == "aa_gram":
        return sum(24 * M * min(e + 1, m1) for k in its for e in range(k + 1)) if depth > 0 else 0
    return 0


KERNEL_OF = {"z": "k_zline<Z_UPD>", "z_aa": "k_zline<Z_AA>", "z_cg": "k_zline<Z_CG>", "y": "k_ypass_r",
             "x": "k_xpass_h", "aa_gram": "k_aa_gram"}


def measured_traffic(kind, n, method):
    """HBM bytes per launch (read + write) of a kernel kind from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (profiles/traffic.txt), or None."""
    path = os.path.join(ROOT, "profiles", "traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 5 and not line.startswith("#") and f[0] == kind and f[1] == str(n) and f[2] == method:
            return float(f[3]) + float(f[4])
    return None


def cpu_baseline_sample(lam, mu, gs, gpu_iters, timed=2):
    """FP64 oracle (NumPy/SciPy) plain LS on the host cores: one warm-up and
    `timed` timed iterations of one load case, extrapolated to the GPU's
    measured iteration counts of the six load cases."""
    import numpy as np

    from oracle import ls_oracle as O

    workers = len(os.sched_getaffinity(0))
    lam64 = lam.astype(np.float64)
    mu64 = mu.astype(np.float64)
    e = np.zeros(6)
    e[0] = 1.0
    O.lippmann_schwinger_rfft(lam64, mu64, e, gs, rtol=0.0, depth=0, maxiter=1, workers=workers)
    t0 = time.perf_counter()
    O.lippmann_schwinger_rfft(lam64, mu64, e, gs, rtol=0.0, depth=0, maxiter=timed, workers=workers)
    t1 = time.perf_counter()
    # one call = setup residual (1 FFT) + `timed` iterations; subtract the setup share
    s_per_it = (t1 - t0) / (timed + 0.5)
    total_its = int(sum(gpu_iters))
    s_per_solve = s_per_it * total_its
    cpu = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {
        "value": 1.0 / s_per_solve,
        "unit": "solves/s",
        "cores": workers,
        "kind": "port",
        "sample": (f"FP64 NumPy/SciPy restatement of the reference solver, plain LS, one 256^3 load case: "
                   f"{timed} timed iterations ({s_per_it:.2f} s/iteration, {workers} threads, {cpu}), "
                   f"extrapolated to the {total_its} iterations the GPU needed for the six load cases"),
        "s_per_iteration": s_per_it,
        "extrapolated": True,
    }


BATCH_METRIC = ("C_eff solves/sec: uneven batch of 384 seeded cross-ply microstructures on grids of "
                "32^3-128^3, six load cases each, to rtol 1e-6")


def batch_record(depth=4, cg=False, count=384):
    """The uneven family of 384 seeded microstructures (grids uniform over
    32^3, 48^3, 64^3, 96^3, 128^3; jaxmaterials_amd.data.batch) solved grid
    group by grid group, concurrently on their own streams, with graph replay
    and autotuned tiles.  The time runs from the moment every grid has been
    tuned (generation of the later microstructures included).  Also returns
    the modelled HBM bytes of the LS passes of the batch (model_bytes per
    microstructure from its six iteration counts, rebases not counted)."""
    import numpy as np

    from jaxmaterials_amd.data.batch import sample_params, solve_batch

    params = [sample_params(i) for i in range(count)]
    res = solve_batch(params, depth=depth, graph=True, autotune=True, concurrent=True, cg=cg)
    its = res["iters"]
    kinds = ("z_cg", "x", "y") if cg else ("z_aa", "x", "y", "aa_gram")
    nbytes = 0.0
    for j, p in enumerate(params):
        n = p["grid"]
        N, M = n ** 3, n * n * (n // 2 + 1)
        nbytes += sum(model_bytes(k, N, M, its[j], 0 if cg else depth, 0, cg) for k in kinds)
    return {"microstructures": count, "method": "cg" if cg else f"aa{depth}", "wall_s": res["wall_s"],
            "ceff_solves_per_s": count / res["wall_s"], "iterations_mean": float(np.mean(its)),
            "model_bytes": nbytes, "achieved_GBps": nbytes / res["wall_s"] / 1e9,
            "grids": {str(k): {kk: vv for kk, vv in v.items() if kk != "autotune"} for k, v in res["timing"].items()},
            "params": params, "iters": its}


def batch_cpu_baseline(rec, timed=2):
    """FP64 oracle conjugate gradients on the host cores: for the first
    microstructure of each grid size, one warm-up and `timed` timed iterations
    of the load case e_0; the batch total extrapolates the seconds per
    iteration of each grid to the GPU's iteration counts of all its load cases."""
    import numpy as np

    from jaxmaterials_amd.common import GridSpec
    from jaxmaterials_amd.data.batch import lame_fields
    from oracle import ls_oracle as O

    workers = len(os.sched_getaffinity(0))
    params, its = rec["params"], rec["iters"]
    s_it = {}
    for j, p in enumerate(params):
        n = p["grid"]
        if n in s_it:
            continue
        lam, mu, _, _ = lame_fields(p, device="cuda")
        lam64, mu64 = lam.cpu().numpy().astype(np.float64), mu.cpu().numpy().astype(np.float64)
        gs = GridSpec(n, n, n, float(n), float(n), float(n))
        e = np.zeros(6)
        e[0] = 1.0
        O.lippmann_schwinger_cg(lam64, mu64, e, gs, rtol=0.0, maxiter=1, workers=workers)
        t0 = time.perf_counter()
        O.lippmann_schwinger_cg(lam64, mu64, e, gs, rtol=0.0, maxiter=timed, workers=workers)
        s_it[n] = (time.perf_counter() - t0) / timed
    total = sum(s_it[p["grid"]] * float(np.sum(its[j])) for j, p in enumerate(params))
    return {"value": len(params) / total, "unit": "solves/s", "cores": workers, "kind": "port",
            "sample": (f"FP64 NumPy/SciPy oracle conjugate gradients, load case e_0 of the first microstructure of "
                       f"each grid, {timed} timed iterations each (s/iteration "
                       + ", ".join(f"{n}^3 {v:.3g}" for n, v in sorted(s_it.items()))
                       + f"), extrapolated to the GPU's {int(np.sum(its))} load-case iterations of the batch"),
            "extrapolated": True}


def _strip_batch(rec):
    """The record without the per-sample parameters and iteration counts."""
    rec.pop("params", None)
    rec.pop("iters", None)
    return rec


def batch_main(args):
    recs = [batch_record(depth=4), batch_record(cg=True)]
    best = max(recs, key=lambda r: r["ceff_solves_per_s"])
    cpu = None if args.no_cpu_baseline else batch_cpu_baseline(best)
    roof = {"bound": "hbm", "kernel": "all LS passes of the batch (aggregate over the concurrent grid groups)",
            "achieved": best["achieved_GBps"], "peak": 8000.0, "unit": "GB/s",
            "frac": best["achieved_GBps"] / 8000.0, "traffic": None,
            "algorithmic_bytes": best["model_bytes"],
            "note": ("modelled LS-pass bytes of the whole batch over its wall time (no per-kernel events: five "
                     "streams run at once); one call's working set (spectrum of 6 load cases x microstructures "
                     "per call, 0.4-3.7 GB) exceeds the 256 MB Infinity Cache, except the Lame fields and "
                     "small per-call state")}
    for r in recs:
        _strip_batch(r)
    line = {"metric": BATCH_METRIC, "value": best["ceff_solves_per_s"], "unit": "solves/s", "n_gpus": 1,
            "steps": 1, "warmup": 0, "ms_per_step": best["wall_s"] * 1e3, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
            "config": {"workload": "uneven batch of 384 seeded cross-ply microstructures, grids 32^3-128^3, layers "
                                   "2-8, radius 1.5-5, vf 0.3-0.6, mu ratio 5-100, six load cases each, rtol 1e-6",
                       "method": best["method"]},
            "methods": {r["method"]: r for r in recs}, "roofline": roof, "cpu_baseline": cpu}
    print(json.dumps(line))


DATASET_METRIC = ("dataset samples/sec: reference layered-fibre generator at 160x160x64 "
                  "(placement + rasterisation + C_eff at rtol = atol = 1e-4)")


def _dataset_generator(gs, n_samples, seed):
    """The reference's dataset script (src/generate_data.py:14-55): 0.5 x 0.5 x 0.2
    domain, 3 fibre layers between 0.01-thick void slabs, volume fraction 0.3,
    radii from the clipped normal law (7.5e-3 +- 0.5e-3 in [5e-3, 1e-2])."""
    import numpy as np

    from jaxmaterials_amd.data.fibres import FibreRadiusDistribution, LayeredFibresDatasetGenerator

    rd = FibreRadiusDistribution(r_avg=7.5e-3, r_min=5.0e-3, r_max=10.0e-3, sigma=0.5e-3, gaussian=True)
    return LayeredFibresDatasetGenerator(gs, 3, 0.01, n_samples, rd, volume_fraction=0.3, mu_fibre=0.5,
                                         mu_material=1.0, mu_void=0.1, lambda_fibre=0.1, lambda_material=0.5,
                                         lambda_void=1.0, dtype=np.float32, rng=np.random.default_rng(seed))


def dataset_main(args):
    """One step = generate() of --samples samples of the reference's dataset
    workload: placement on the host, GPU rasterisation, all load cases of up
    to --batch samples per plain-LS solver call (the reference's iteration),
    fields and tensors copied to host arrays as the reference's generate()
    leaves them."""
    os.environ["LS_TIMING"] = "1"
    import numpy as np
    import torch

    from jaxmaterials_amd.common import GridSpec

    gs = GridSpec(160, 160, 64, 0.5, 0.5, 0.2)
    N = gs.nx * gs.ny * gs.nz
    M = gs.nx * gs.ny * (gs.nz // 2 + 1)
    for w in range(args.warmup):
        _dataset_generator(gs, min(args.samples, args.batch), 1000 + w).generate(max_batch=args.batch)
    times, gens = [], []
    for step in range(args.steps):
        gen = _dataset_generator(gs, args.samples, 5713853 + step)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        gen.generate(max_batch=args.batch)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        gens.append(gen)
    sec = float(np.mean(times))
    total = {k: float(np.mean([g.timing[k] for g in gens])) for k in ("placement_s", "rasterise_s", "solve_s",
                                                                       "copy_s")}
    # algorithmic bytes of the solver passes, batch by batch (padding slots repeat the last sample)
    kinds = {}
    k = gens[0].timing["per_call"]
    for name in ("z", "y", "x"):
        b = 0
        for g in gens:
            its = g.iterations
            for c0 in range(0, len(its), k):
                chunk = [its[min(c0 + q, len(its) - 1)] for q in range(k)]
                flat = [int(i) for row in chunk for i in row]
                b += model_bytes(name, N, M, flat, 0, 0) + (
                    sum(max(int(i) for i in row) for row in chunk) - max(flat)) * 8 * N * (name == "z")
        ms = sum(g.timing["solver_ms"].get(name, 0.0) for g in gens)
        launches = sum(g.timing["solver_launches"].get(name, 0) for g in gens)
        kinds[name] = {"ms_total": ms / len(gens), "launches": launches // len(gens),
                       "algorithmic_bytes": b / len(gens), "achieved_GBps": b / (ms / 1e3) / 1e9 if ms else None}
    dom = max(kinds, key=lambda q: kinds[q]["ms_total"])
    dk = kinds[dom]
    iters = np.concatenate([g.iterations.reshape(-1) for g in gens])
    cpu = None
    if not args.no_cpu_baseline:
        cpu = dataset_cpu_baseline(gs, gens[0], float(np.mean(iters)))
    line = {
        "metric": DATASET_METRIC,
        "value": args.samples / sec,
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"reference dataset script: 160x160x64 grid of 0.5 x 0.5 x 0.2, 3 layers, "
                               f"void 0.01, vf 0.3, radii 7.5e-3 +- 0.5e-3; {args.samples} samples per step, "
                               f"six load cases each, plain LS to rtol = atol = 1e-4, maxiter 100, "
                               f"{k} samples per solver call",
                   "grid": [160, 160, 64], "samples_per_step": args.samples},
        "phases_s_per_step": total,
        "iterations_mean": float(np.mean(iters)),
        "roofline": {"bound": "hbm", "kernel": {"z": "k_zline<Z_UPD>", "y": "k_ypass", "x": "k_xpass"}[dom],
                     "achieved": dk["achieved_GBps"], "peak": 8000.0, "unit": "GB/s",
                     "frac": dk["achieved_GBps"] / 8000.0 if dk["achieved_GBps"] else None, "traffic": None,
                     "passes": kinds},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


def dataset_cpu_baseline(gs, gen, mean_iters):
    """The reference's host cost of one sample: its NumPy rasteriser (the
    oracle's literal restatement of data.py:267-314) timed on one sample, plus
    the FP64 oracle plain LS timed for two iterations of one load case and
    extrapolated to six load cases at the GPU's mean iteration count.  The
    placement (about 10 ms of compiled code per sample) is left out."""
    import numpy as np

    from oracle import fibres_oracle as fo
    from oracle import ls_oracle as O

    workers = len(os.sched_getaffinity(0))
    ref = _dataset_generator(gs, 1, 5713853)
    _, pos, rad, ori = ref.generate_fibre_positions()
    t0 = time.perf_counter()
    f = fo.material_properties(gs, ref.d_void, pos, rad, ori, (ref.mu_void, ref.mu_fibre, ref.mu_material),
                               (ref.lambda_void, ref.lambda_fibre, ref.lambda_material), dtype=np.float32)
    t_raster = time.perf_counter() - t0
    lam64, mu64 = f[0].astype(np.float64), f[1].astype(np.float64)
    e = np.zeros(6)
    e[0] = 1.0
    O.lippmann_schwinger_rfft(lam64, mu64, e, gs, rtol=0.0, depth=0, maxiter=1, workers=workers)
    t0 = time.perf_counter()
    O.lippmann_schwinger_rfft(lam64, mu64, e, gs, rtol=0.0, depth=0, maxiter=2, workers=workers)
    s_it = (time.perf_counter() - t0) / 2.5
    s_sample = t_raster + 6 * mean_iters * s_it
    return {"value": 1.0 / s_sample, "unit": "samples/s", "cores": workers, "kind": "port",
            "sample": (f"one sample: NumPy rasteriser {t_raster:.2f} s (1 thread) + FP64 NumPy/SciPy plain LS "
                       f"{s_it:.3f} s/iteration ({workers} threads) x 6 load cases x {mean_iters:.1f} iterations "
                       f"(extrapolated from 2 timed iterations)"),
            "rasterise_s": t_raster, "s_per_iteration": s_it, "extrapolated": True}


PHASEFIELD_METRIC = ("phase-field staggered load steps/sec: the reference's phase-field script "
                     "(55x50x1 notched plate, 51 load steps, AA(3) elasticity to 1e-6, FP64 phase field to 1e-9)")


def _notched_plate(gs, lam_solid=121.15, mu_solid=80.77, margin=0.05, crack_length=0.5):
    """Lame fields of the reference's phase-field script (tests/phasefield.py:
    45-131 with shape="homogeneous"): a homogeneous plate, zero stiffness in the
    margins x < margin and x > Lx - margin, and a one-voxel crack from the
    left margin along y = Ly / 2.  Returns float32 (lmbda, mu)."""
    import numpy as np

    h = [gs.Lx / gs.nx, gs.Ly / gs.ny, gs.Lz / gs.nz]
    X, Y, _ = np.meshgrid(*[hh * (0.5 + np.arange(n)) for n, hh in zip((gs.nx, gs.ny, gs.nz), h)], indexing="ij")
    lam = np.full(X.shape, lam_solid)
    mu = np.full(X.shape, mu_solid)
    margin_bound = (X < margin) | (X > gs.Lx - margin)
    crack = (X <= margin + crack_length) & (Y > gs.Ly / 2 - h[1]) & (Y < gs.Ly / 2)
    lam[margin_bound | crack] = 0.0
    mu[margin_bound | crack] = 0.0
    return lam.astype(np.float32), mu.astype(np.float32)


def phasefield_main(args):
    """One step = one staggered load step of the reference's phase-field script
    (tests/phasefield.py:160-225): phase field (FP64 fixed point to 1e-9), the
    damaged elasticity solve (Anderson depth 3 to rtol 1e-6), the history
    update; eps_yy from 0 to 0.01 in 51 steps."""
    import numpy as np
    import torch

    from jaxmaterials_amd.common import GridSpec
    from jaxmaterials_amd.solver.phasefield import staggered_steps

    gs = GridSpec(55, 50, 1, 1.0, 1.0, 0.015)
    lam, mu = _notched_plate(gs)
    ebars = [np.array([0.0, e, 0.0, 0.0, 0.0, 0.0]) for e in np.linspace(0, 0.01, 51)]
    kw = dict(gc=2.7e-3, lc=0.015, k_stab=1e-6, depth=3, rtol=1e-6, maxiter=1000, pf_tolerance=1e-9, pf_maxiter=10000)
    lt, mt = torch.from_numpy(lam).cuda(), torch.from_numpy(mu).cuda()
    for _ in range(args.warmup):
        staggered_steps(lt, mt, ebars[:3], gs, **kw)
    times, out = [], None
    for _ in range(args.steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = staggered_steps(lt, mt, ebars, gs, **kw)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    sec = float(np.mean(times))
    its = [st["elasticity_iterations"] for st in out["steps"]]
    pfits = [st["pf_iterations"] for st in out["steps"]]
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import phasefield_oracle as PF

        t0 = time.perf_counter()
        d = np.zeros(lam.shape)
        H = np.zeros(lam.shape)
        sig_o = None
        for eb in ebars:
            d, _ = PF.phase_field_solve(H, d, kw["gc"], kw["lc"], gs, tolerance=1e-9, maxiter=10000)
            if np.any(eb):
                eps, sig_o, _ = PF.elasticity_solve(lam, mu, eb, d.astype(np.float32), kw["k_stab"], gs, rtol=1e-6,
                                                    depth=3, maxiter=1000)
                H = np.maximum(H, PF.strain_energy_plus(lam, mu, eps))
        t_cpu = time.perf_counter() - t0
        cpu = {"value": len(ebars) / t_cpu, "unit": "steps/s", "cores": 1, "kind": "port",
               "sample": f"the whole 51-step script on the FP64 NumPy restatement, {t_cpu:.1f} s (NumPy FFTs, "
                         f"one thread)",
               "final_sigma_yy": float(sig_o.mean(axis=(1, 2, 3))[1]) if sig_o is not None else None}
    line = {
        "metric": PHASEFIELD_METRIC,
        "value": len(ebars) / sec,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 elasticity, fp64 phase field",
        "data": "synthetic",
        "config": {"workload": "reference phase-field script: 55x50x1 plate of 1 x 1 x 0.015, lambda 121.15, "
                               "mu 80.77, zero stiffness in 0.05 margins and a 0.5 crack, eps_yy 0 -> 0.01 in 51 "
                               "steps, gc 2.7e-3, lc 0.015, k 1e-6",
                   "grid": [55, 50, 1], "load_steps": len(ebars)},
        "script_seconds": sec,
        "elasticity_iterations_total": int(sum(its)),
        "phase_field_iterations_total": int(sum(pfits)),
        "final_sigma_yy": out["steps"][-1]["sigma_mean"][1],
        "max_damage": float(out["damage"].max()),
        "roofline": None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--depth", type=int, default=6)
    ap.add_argument("--methods", default="plain,aa,aa5,aa_ws,cg,cg_ws",
                    help="plain: no Anderson; aa: Anderson of --depth; aaK: Anderson of depth K (aa5 is the "
                         "configuration's m = 5); cg: conjugate gradients; a suffix _ws starts from the learned "
                         "warm start")
    ap.add_argument("--rtol", type=float, default=1e-6)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", choices=("crossply256", "batch384", "dataset", "phasefield"), default="crossply256",
                    help="crossply256: the C_eff solve benchmark (default); batch384: the uneven batch of 384 "
                         "microstructures; dataset: the reference's layered-fibre dataset generation; phasefield: "
                         "the reference's phase-field script")
    ap.add_argument("--no-batch", action="store_true", help="crossply256: leave out the 384-microstructure batch")
    ap.add_argument("--samples", type=int, default=128, help="dataset workload: samples per step")
    ap.add_argument("--batch", type=int, default=32, help="dataset workload: samples per solver call")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the solver runs each solve on a single MI355X and this build has no "
              "multi-GPU path (--gpus must be 1).", file=sys.stderr)
        sys.exit(2)
    if args.workload == "dataset":
        dataset_main(args)
        return
    if args.workload == "phasefield":
        phasefield_main(args)
        return
    if args.workload == "batch384":
        batch_main(args)
        return

    import numpy as np
    import torch

    from jaxmaterials_amd.common import GridSpec
    from jaxmaterials_amd.data import crossply_lame
    from jaxmaterials_amd.solver.device import LSSolver, unit_load_cases

    n = args.grid
    layers = 16 if n == 256 else max(2, n // 16)
    radius = 6.0 if n == 256 else 6.0 * n / 256
    shape = (n, n, n)
    ext = (float(n),) * 3
    gs = GridSpec(*shape, *ext)
    N = n**3
    M = n * n * (n // 2 + 1)
    lam, mu, vf, nfib = crossply_lame(shape, ext, layers, radius, 0.5, seed=2)
    print(f"# microstructure {n}^3: {nfib} fibres, achieved vf {vf:.4f}", file=sys.stderr, flush=True)

    golden = None
    gpath = os.path.join(ROOT, "tests", "golden", "crossply256_ceff.json")
    if n == 256 and os.path.exists(gpath):
        with open(gpath) as f:
            golden = np.array(json.load(f)["ceff"])

    results = {}
    net = None
    for method in args.methods.split(","):
        warm = method.endswith("_ws")
        base_name = method[:-3] if warm else method
        use_cg = base_name == "cg"
        if base_name in ("plain", "cg"):
            depth = 0
        elif base_name == "aa":
            depth = args.depth
        elif base_name.startswith("aa") and base_name[2:].isdigit():
            depth = int(base_name[2:])
        else:
            raise SystemExit(f"unknown method {method}")
        solver = LSSolver(gs, nrhs=6, depth=depth)
        loads = unit_load_cases()
        eps = torch.empty((6, 6) + shape, dtype=torch.float32, device="cuda")
        ws_info = None
        if warm:
            from jaxmaterials_amd.warmstart import WarmStartNet
            from jaxmaterials_amd.warmstart.train import DEFAULT_WEIGHTS, warm_start_solve
            if net is None:
                net = WarmStartNet.load(DEFAULT_WEIGHTS)
            st0 = warm_start_solve(net, solver, lam, mu, loads, epsilon=eps, rtol=args.rtol, maxiter=0,
                                   check=False, cg=use_cg)
            r_init = st0["residual"]
            # equilibrium residual norm relative to the uniform start's (R divides by each start's mean stress)
            la, mm = float(lam.double().mean()), float(mu.double().mean())
            m0 = np.array([[la + 2 * mm * (c == k) if c < 3 and k < 3 else (2 * mm if c == k else 0.0)
                            for c in range(6)] for k in range(6)])
            wv = np.array([1, 1, 1, 2, 2, 2.0])
            a_init = r_init * np.sqrt((wv * st0["sigma_mean"] ** 2).sum(1)) / np.sqrt((wv * m0 ** 2).sum(1))
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            net.forward(lam.reshape((1,) + shape), mu.reshape((1,) + shape), eps, extents=ext)
            torch.cuda.synchronize()
            ws_info = {"initial_residual_vs_uniform": [float(v) for v in r_init],
                       "initial_equilibrium_residual_norm_vs_uniform": [float(v) for v in a_init],
                       "inference_ms": 1e3 * (time.perf_counter() - t0),
                       "weights": os.path.basename(DEFAULT_WEIGHTS)}

        def run(solver=solver, eps=eps, warm=warm, use_cg=use_cg):
            # the iterations between host synchronisations replayed as captured HIP graphs
            if warm:
                return warm_start_solve(net, solver, lam, mu, loads, epsilon=eps, rtol=args.rtol, maxiter=3000,
                                        cg=use_cg, graph=True)
            return solver.solve(lam, mu, loads, epsilon=eps, rtol=args.rtol, maxiter=3000, cg=use_cg, graph=True)

        def timed(k, run=run):
            ts, o = [], None
            for _ in range(k):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                o = run()
                torch.cuda.synchronize()
                ts.append(time.perf_counter() - t0)
            return ts, o

        for _ in range(args.warmup):
            run()
        # every method is timed for up to 5 solves; the fastest again for --steps below
        times, out = timed(min(args.steps, 5))
        # per-pass kernel times from one more solve, launch by launch with HIP events on the solver stream
        os.environ["LS_TIMING"] = "1"
        run()
        del os.environ["LS_TIMING"]
        ms_kind, cnt_kind, nreb = solver.last_timing()
        iters = [int(i) for i in out["iters"]]
        C = out["sigma_mean"]
        sec = float(np.mean(times))
        kinds = {}
        for name in ("z", "z_aa", "z_cg", "y", "x", "aa_gram"):
            if cnt_kind[name] == 0:
                continue
            b = model_bytes(name, N, M, iters, depth, nreb, use_cg)
            t = ms_kind[name] / 1e3
            kinds[name] = {"kernel": KERNEL_OF[name], "ms_total": t * 1e3, "launches": int(cnt_kind[name]),
                           "avg_launch_ms": t * 1e3 / max(1, cnt_kind[name]),
                           "algorithmic_bytes": b, "achieved_GBps": (b / t / 1e9) if t > 0 else None}
        loop_bytes = sum(model_bytes(k, N, M, iters, depth, nreb, use_cg) for k in kinds)
        rec = {"solves_per_s": 1.0 / sec, "ms_per_solve": sec * 1e3, "timed_solves": len(times), "iters": iters,
               "depth": depth, "accelerator": "cg" if use_cg else ("anderson" if depth > 0 else "none"),
               "base_reevaluations": nreb, "passes": kinds, "loop_GBps": loop_bytes / sec / 1e9,
               "kernel_ms_per_solve": sum(ms_kind.values()), "ceff": C.tolist(), "_run": run, "_timed": timed,
               "_loop_bytes": loop_bytes}
        if ws_info is not None:
            rec["warm_start"] = ws_info
        if golden is not None:
            sig = np.abs(golden) > 1e-3 * np.abs(golden).max()
            rec["ceff_max_rel_err_vs_fp64_oracle"] = float((np.abs(C - golden)[sig] / np.abs(golden)[sig]).max())
        results[method] = rec
        print(f"# {method}: {rec['ms_per_solve']:.1f} ms/solve, iters {iters}, loop {rec['loop_GBps']:.0f} GB/s, "
              + ", ".join(f"{k} {v['achieved_GBps']:.0f} GB/s ({v['ms_total']:.0f} ms)" for k, v in kinds.items()
                          if v['achieved_GBps']), file=sys.stderr, flush=True)
        rec["_solver"], rec["_eps"] = solver, eps
        del solver, eps

    best_name = max(results, key=lambda k: results[k]["solves_per_s"])
    best = results[best_name]
    # the fastest method timed for --steps solves (the headline value)
    if args.steps > best["timed_solves"]:
        times, _ = best["_timed"](args.steps)
        sec = float(np.mean(times))
        best.update(solves_per_s=1.0 / sec, ms_per_solve=sec * 1e3, timed_solves=len(times),
                    loop_GBps=best["_loop_bytes"] / sec / 1e9)
    for k in list(results):
        results[k]["_solver"].close()
        for key in [q for q in results[k] if q.startswith("_")]:
            del results[k][key]
    if net is not None:
        net.release()
    torch.cuda.empty_cache()
    dom = max(best["passes"], key=lambda k: best["passes"][k]["ms_total"])
    dk = best["passes"][dom]
    peak = 8000.0
    roofline = {"bound": "hbm", "kernel": dk["kernel"], "achieved": dk["achieved_GBps"], "peak": peak,
                "unit": "GB/s", "frac": dk["achieved_GBps"] / peak,
                "traffic": measured_traffic(dom, n, best_name.replace("_ws", "")),
                "algorithmic_bytes_per_launch": dk["algorithmic_bytes"] / max(1, dk["launches"]),
                "avg_launch_ms": dk["avg_launch_ms"],
                "loop_achieved": best["loop_GBps"], "loop_frac": best["loop_GBps"] / peak}
    # measured streaming ceilings of this GPU (grid-stride float4 kernels)
    import ctypes

    from jaxmaterials_amd import _lib

    ceil = np.zeros(4)
    torch.cuda.synchronize()
    if _lib.load().ls_hbm_stream_probe(4 << 30, 5, ceil.ctypes.data_as(ctypes.c_void_p), None) == 0:
        roofline["stream_ceiling_GBps"] = {"copy": ceil[0], "read": ceil[1], "write": ceil[2],
                                           "read2_write1": ceil[3]}
        roofline["frac_of_read2_write1_ceiling"] = dk["achieved_GBps"] / ceil[3]
    cpu = None
    if not args.no_cpu_baseline and "plain" in results:
        cpu = cpu_baseline_sample(lam.cpu().numpy(), mu.cpu().numpy(), gs, results["plain"]["iters"])
    line = {
        "metric": METRIC,
        "value": best["solves_per_s"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": best["ms_per_solve"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"crossply {n}^3, {layers} layers, fibre radius {radius:g} vox, RSA vf 0.5 "
                               f"(achieved {vf:.3f}), 6 unit load cases batched, rtol {args.rtol:g}",
                   "method": best_name, "anderson_depth": best["depth"], "grid": n,
                   "reductions": "fp64"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "methods": {k: {kk: vv for kk, vv in v.items() if kk != "ceff"} for k, v in results.items()},
    }
    if not args.no_batch and n == 256:
        line["batch384"] = {"aa4": _strip_batch(batch_record(depth=4)), "cg": _strip_batch(batch_record(cg=True))}
    if "plain" in results and len(results) > 1:
        line["speedup_vs_plain"] = {k: v["solves_per_s"] / results["plain"]["solves_per_s"]
                                    for k, v in results.items() if k != "plain"}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
