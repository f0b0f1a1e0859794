"""Training of the warm-start network on solver-generated pairs, and the
warm-started effective-tensor solve.

Samples come from the 256^3 benchmark family scaled to 64^3: cross-ply fibre
microstructures with layers of 16 voxels, fibre radius 6 voxels, RSA volume
fraction 0.5, matrix (lambda, mu) = (2.593, 1.111), fibre (19.44, 29.17),
each solved for the six unit load cases to a relative residual of 1e-6 by
the batched conjugate-gradient solver.  Inputs are the Lame fields; targets
the strain fluctuations eps - e_k of the six load cases (bf16, 36 channels).
The network predicts displacements, so every start is compatible; the loss
is the relative strain energy of the error (ls_ws_residual_loss), the norm
conjugate gradients reduce, optionally plus the squared equilibrium residual
relative to the uniform start's (the solver's stop measure).

    python -m jaxmaterials_amd.warmstart.train --samples 4000 --epochs 4 --out weights.npz
"""

import argparse
import math
import os
import sys
import time

import numpy as np
import torch

from ..common import GridSpec
from ..data import crossply_lame
from ..solver.device import LSSolver
from .net import NOUT, WarmStartNet

__all__ = ["generate", "train", "warm_start_solve", "DEFAULT_WEIGHTS"]

DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "crossply64.npz")


def _family(grid):
    """Cross-ply parameters of the benchmark family at grid^3 (16-voxel layers, r = 6)."""
    return max(1, grid // 16), 6.0, 0.5


def generate(net, n_samples, grid=64, seed0=1_000_000, batch=16, rtol=1e-6, log=print):
    """Lame fields [S][g][g][g] and bf16 targets [S][36][g][g][g] (device)."""
    dev = net.device
    shape = (grid,) * 3
    ext = (float(grid),) * 3
    gs = GridSpec(*shape, *ext)
    layers, radius, vf = _family(grid)
    lam = torch.empty((n_samples,) + shape, dtype=torch.float32, device=dev)
    mu = torch.empty_like(lam)
    tgt = torch.empty((n_samples, NOUT) + shape, dtype=torch.int16, device=dev)
    solver = LSSolver(gs, nrhs=6 * batch, nmat=batch)
    eb_host = np.tile(np.eye(6, dtype=np.float32), (batch, 1))
    eb_dev = torch.tensor(eb_host, device=dev)
    mats = np.repeat(np.arange(batch, dtype=np.int32), 6)
    lb = torch.empty((batch,) + shape, dtype=torch.float32, device=dev)
    mb = torch.empty_like(lb)
    t0 = time.perf_counter()
    its = []
    for s0 in range(0, n_samples, batch):
        nb = min(batch, n_samples - s0)
        for j in range(batch):
            l, m, _, _ = crossply_lame(shape, ext, layers, radius, vf, seed0 + s0 + min(j, nb - 1), device=dev)
            lb[j].copy_(l)
            mb[j].copy_(m)
        out = solver.solve(lb, mb, eb_host, rhs_mat=mats, rtol=rtol, maxiter=3000, cg=True)
        its.append(out["iters"][: 6 * nb])
        lam[s0:s0 + nb].copy_(lb[:nb])
        mu[s0:s0 + nb].copy_(mb[:nb])
        net.make_target(out["epsilon"][: 6 * nb], eb_dev[: 6 * nb], out=tgt[s0:s0 + nb].view((6 * nb, 6) + shape))
        if log and (s0 // batch) % 25 == 0:
            log(f"generate: {s0 + nb}/{n_samples} samples, {time.perf_counter() - t0:.1f} s, "
                f"mean iterations {np.mean(its[-1]):.1f}")
    torch.cuda.synchronize(dev)
    solver.close()
    return lam, mu, tgt


def evaluate_loss(net, lam, mu, tgt, idx, sps, extents):
    """Validation metrics over the samples idx: the relative L2 error of the
    fluctuation, the mean relative strain energy of the error and the mean
    relative squared equilibrium residual."""
    shape = tuple(lam.shape[1:])
    pred = torch.empty((sps, NOUT) + shape, dtype=torch.float32, device=net.device)
    num = den = 0.0
    res, en = [], []
    for i in range(0, len(idx) - sps + 1, sps):
        sel = idx[i:i + sps]
        ls, ms = lam.index_select(0, sel).contiguous(), mu.index_select(0, sel).contiguous()
        net.forward(ls, ms, pred, add_eps_bar=False, extents=extents)
        t = tgt.index_select(0, sel).contiguous()
        tf = t.view(torch.bfloat16).float()
        num += float(((pred - tf) ** 2).sum())
        den += float((tf ** 2).sum())
        res.append(float(net.loss_backward(pred, t, ls, ms, wm=0.0, wr=1.0, backward=False).item()))
        en.append(float(net.loss_backward(pred, t, ls, ms, wm=1.0, wr=0.0, backward=False).item()))
    return math.sqrt(num / max(den, 1e-30)), float(np.mean(en)), float(np.mean(res))


def train(net, lam, mu, tgt, epochs=4, sps=4, lr=3e-4, lr_min=1e-5, n_val=64, seed=0, wm=1.0, wr=0.0, log=print,
          time_limit=None, spike=20.0):
    """Adam with linear warm-up and a cosine learning-rate schedule; sps samples per
    step.  A step whose loss is not finite or exceeds `spike` times the running
    mean of the accepted losses is skipped (the residual term weighs the
    displacement's second differences, which makes the problem stiff).
    Returns a dict of losses and timings."""
    dev = net.device
    S = lam.shape[0]
    shape = tuple(lam.shape[1:])
    extents = tuple(float(n) for n in shape)
    n_train = S - n_val
    g = torch.Generator(device="cpu").manual_seed(seed)
    pred = torch.empty((sps, NOUT) + shape, dtype=torch.float32, device=dev)
    steps_per_epoch = n_train // sps
    total = epochs * steps_per_epoch
    val_idx = torch.arange(n_train, S, device=dev)
    e0, n0, r0 = evaluate_loss(net, lam, mu, tgt, val_idx, sps, extents)
    hist = {"val_rel_err": [e0], "val_rel_energy": [n0], "val_rel_residual2": [r0], "train_loss": []}
    log(f"train: {n_train} samples, {total} steps; initial validation relative error {e0:.4f}, "
        f"relative energy {n0:.4e}, relative squared residual {r0:.4e}")
    t0 = time.perf_counter()
    step = 0
    ema = None
    skipped = 0
    for ep in range(epochs):
        perm = torch.randperm(n_train, generator=g).to(dev)
        for i in range(steps_per_epoch):
            sel = perm[i * sps:(i + 1) * sps]
            ls, ms = lam.index_select(0, sel).contiguous(), mu.index_select(0, sel).contiguous()
            net.forward(ls, ms, pred, add_eps_bar=False, extents=extents)
            loss = net.loss_backward(pred, tgt.index_select(0, sel).contiguous(), ls, ms, wm=wm, wr=wr)
            warm = min(1.0, (step + 1) / max(1, total // 50))  # linear warm-up over 2% of the steps
            lr_t = warm * (lr_min + 0.5 * (lr - lr_min) * (1.0 + math.cos(math.pi * step / max(1, total))))
            lv = float(loss.item())
            step += 1
            if not math.isfinite(lv) or (ema is not None and lv > spike * ema):
                skipped += 1
                continue
            ema = lv if ema is None else 0.98 * ema + 0.02 * lv
            net.adam_step_(lr_t)
            if i % 100 == 99:
                hist["train_loss"].append(lv)
                log(f"epoch {ep} step {i + 1}/{steps_per_epoch}: loss {hist['train_loss'][-1]:.4e}, lr {lr_t:.2e}, "
                    f"{time.perf_counter() - t0:.1f} s")
            if time_limit and time.perf_counter() - t0 > time_limit:
                break
        e, en, r = evaluate_loss(net, lam, mu, tgt, val_idx, sps, extents)
        hist["val_rel_err"].append(e)
        hist["val_rel_energy"].append(en)
        hist["val_rel_residual2"].append(r)
        log(f"epoch {ep}: validation relative error {e:.4f}, relative energy {en:.4e}, relative squared residual "
            f"{r:.4e}")
        if time_limit and time.perf_counter() - t0 > time_limit:
            log("train: time limit reached")
            break
    torch.cuda.synchronize(dev)
    hist["seconds"] = time.perf_counter() - t0
    hist["steps"] = step
    hist["skipped"] = skipped
    log(f"train: {skipped} of {step} steps skipped by the spike guard")
    return hist


def warm_start_solve(net, solver, lmbda, mu, epsilon_bar, epsilon=None, **solve_kw):
    """Solve the six unit load cases of one microstructure from the network's
    prediction: one evaluation writes e_k + the predicted fluctuation of every
    load case into the solver's strain buffer (e_k + sym(D^+ u_k), compatible
    with mean e_k by construction, so the solver skips its projection); the
    solver falls back to the uniform start where its residual is larger.
    epsilon_bar must be the six unit load cases."""
    shape = solver.shape
    if solver.nrhs != 6 or not np.array_equal(np.asarray(epsilon_bar, dtype=np.float32).reshape(-1, 6),
                                              np.eye(6, dtype=np.float32)):
        raise ValueError("the warm start predicts the six unit load cases e_0..e_5 of one microstructure")
    if epsilon is None:
        epsilon = torch.empty((6, 6) + shape, dtype=torch.float32, device=net.device)
    net.forward(lmbda.reshape((1,) + shape), mu.reshape((1,) + shape), epsilon, add_eps_bar=True,
                extents=(solver.gs.Lx, solver.gs.Ly, solver.gs.Lz))
    torch.cuda.synchronize(net.device)
    return solver.solve(lmbda, mu, epsilon_bar, epsilon=epsilon, warm_start=True, compatible_start=True, **solve_kw)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--samples", type=int, default=4000)
    ap.add_argument("--grid", type=int, default=64)
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--sps", type=int, default=4)
    ap.add_argument("--lr", type=float, default=3e-4)
    ap.add_argument("--time-limit", type=float, default=None)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--wm", type=float, default=1.0, help="weight of the relative strain energy of the error")
    ap.add_argument("--wr", type=float, default=0.0, help="weight of the relative squared equilibrium residual")
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    args = ap.parse_args(argv)
    log = lambda s: print(s, flush=True)
    net = WarmStartNet(seed=args.seed)
    t0 = time.perf_counter()
    lam, mu, tgt = generate(net, args.samples, grid=args.grid, log=log)
    tgen = time.perf_counter() - t0
    log(f"generated {args.samples} samples in {tgen:.1f} s")
    hist = train(net, lam, mu, tgt, epochs=args.epochs, sps=args.sps, lr=args.lr, seed=args.seed, wm=args.wm,
                 wr=args.wr, log=log, time_limit=args.time_limit)
    net.save(args.out, samples=args.samples, grid=args.grid, epochs=args.epochs, steps=hist["steps"],
             val_rel_err=np.array(hist["val_rel_err"]), val_rel_energy=np.array(hist["val_rel_energy"]),
             val_rel_residual2=np.array(hist["val_rel_residual2"]), train_seconds=hist["seconds"],
             generate_seconds=tgen, wm=args.wm, wr=args.wr, lr=args.lr)
    log(f"saved {args.out}: validation relative error {hist['val_rel_err'][-1]:.4f}, "
        f"{hist['steps']} steps in {hist['seconds']:.1f} s")
    return 0


if __name__ == "__main__":
    sys.exit(main())
