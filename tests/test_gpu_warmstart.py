"""GPU parity of the warm-start network kernels (csrc/conv3d.hip, through the C
ABI) against the CPU FP64 oracle of oracle/warmstart_oracle.py, which rounds
to bf16 where the GPU stores bf16.

Tolerances: forward output within 1e-2 relative L2 (bf16 storage of 15
activations, FP32 accumulation); loss within 1e-3 relative; every layer's
weight and bias gradient within 3e-2 relative L2 (the GPU also stores the
backpropagated gradients in bf16, the oracle keeps them in FP64); the
residual-loss kernel alone: value within 1e-5, dL/dpred (bf16) within 1e-2.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bf16_bits(a):
    """float32 array -> bf16 bit patterns (round to nearest even) as int16."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    return r.view(np.int16)


EXT = (16.0, 14.0, 12.0)
# largest residual of the trained network's start relative to the uniform
# start's on held-out samples (every load case)
WS_MAX_RATIO = 0.7


def _setup(gpu, n=16, S=2, seed=3):
    import torch

    from jaxmaterials_amd.warmstart.net import HEAD_VALID, NOUT, WarmStartNet

    rng = np.random.default_rng(seed)
    net = WarmStartNet(device=gpu, seed=seed)
    # random nonzero biases so that the bias paths are exercised
    sd = net.state_dict()
    for k, (w, b) in sd.items():
        b[:] = rng.uniform(-0.05, 0.05, b.shape)
        if k == "head":
            b[HEAD_VALID:] = 0.0
    net.load_state_dict(sd)
    lam = rng.choice([2.593, 19.44], size=(S, n, n, n)).astype(np.float32)
    mu = np.where(lam > 3.0, 29.17, 1.111).astype(np.float32)
    target = (rng.standard_normal((S, NOUT, n, n, n)) * 0.2).astype(np.float32)
    dev = dict(device=gpu)
    t = dict(lam=torch.tensor(lam, **dev), mu=torch.tensor(mu, **dev), target=torch.tensor(_bf16_bits(target), **dev))
    return net, lam, mu, target, t, S


def _oracle(net, lam, mu, target, wm):
    from oracle import warmstart_oracle as W

    sd = {k: (w.astype(np.float64), b.astype(np.float64)) for k, (w, b) in net.state_dict().items()}
    return W.loss_and_grads(sd, lam.astype(np.float64), mu.astype(np.float64), net.norm, target.astype(np.float64),
                            EXT, wm=wm, wr=1.0, bf16=True)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def test_forward_matches_oracle(gpu):
    import torch

    net, lam, mu, target, t, S = _setup(gpu)
    n = lam.shape[-1]
    out = torch.empty((S, 36, n, n, n), dtype=torch.float32, device=gpu)
    net.forward(t["lam"], t["mu"], out, add_eps_bar=True, extents=EXT)
    torch.cuda.synchronize()
    _, pred, _ = _oracle(net, lam, mu, target, 0.1)
    eye = np.eye(6).reshape(1, 36, 1, 1, 1)
    got = out.cpu().numpy()
    err = _rel(got - eye, pred)
    print(f"forward: rel L2 err of the fluctuation {err:.2e}")
    assert err < 1e-2
    assert _rel(got, pred + eye) < 1e-2


def test_residual_loss_kernel_matches_oracle(gpu):
    """ls_ws_residual_loss (loss and dL/dpred) against autograd of the FP64
    oracle residual_loss on random predictions and targets."""
    import ctypes

    import torch

    from jaxmaterials_amd import _lib
    from oracle import warmstart_oracle as W

    rng = np.random.default_rng(11)
    S, n = 2, 16
    lam = rng.choice([2.593, 19.44], size=(S, n, n, n)).astype(np.float32)
    mu = np.where(lam > 3.0, 29.17, 1.111).astype(np.float32)
    target = (rng.standard_normal((S, 36, n, n, n)) * 0.2).astype(np.float32)
    tq = _bf16_bits(target)
    pred = (target + rng.standard_normal(target.shape) * 0.05).astype(np.float32)
    wm = 0.3
    dev = dict(device=gpu)
    pt, tt = torch.tensor(pred, **dev), torch.tensor(tq, **dev)
    lt, mt = torch.tensor(lam, **dev), torch.tensor(mu, **dev)
    grad = torch.zeros((S * n**3, 40), dtype=torch.int16, **dev)
    scratch = torch.empty(54 * S * n**3, dtype=torch.float32, **dev)
    part = torch.empty(24 * S * (n**3 // 256 + 1), dtype=torch.float64, **dev)
    scale = torch.empty(12 * S, dtype=torch.float32, **dev)
    loss = torch.zeros(1, dtype=torch.float64, **dev)
    dims = (ctypes.c_int * 4)(S, n, n, n)
    ext = np.array(EXT, np.float32)
    p = lambda a: ctypes.c_void_p(a.data_ptr())  # noqa: E731
    rc = _lib.load().ls_ws_residual_loss(dims, ext.ctypes.data_as(ctypes.c_void_p), p(pt), p(tt), p(lt), p(mt), wm,
                                         1.0, p(grad), 40, p(scratch), scratch.numel(), p(part), part.numel(), p(scale),
                                         p(loss), None)
    assert rc == 0
    torch.cuda.synchronize()
    tgt64 = torch.tensor(target).to(torch.bfloat16).to(torch.float64)
    pr = torch.tensor(pred.astype(np.float64), requires_grad=True)
    L = W.residual_loss(pr, tgt64, torch.tensor(lam.astype(np.float64)), torch.tensor(mu.astype(np.float64)), EXT, wm,
                        1.0)
    L.backward()
    g = grad.cpu().numpy().view(np.uint16).astype(np.uint32) << 16
    g = g.view(np.float32).reshape(S, n, n, n, 40)
    assert np.all(g[..., 36:] == 0)
    g = np.moveaxis(g[..., :36], -1, 1)
    want = pr.grad.numpy()
    eg = _rel(g, want)
    el = abs(float(loss.item()) - float(L)) / float(L)
    print(f"residual loss: value rel err {el:.2e}, gradient rel L2 {eg:.2e}")
    assert el < 1e-5
    assert eg < 1e-2


def test_gradients_match_oracle(gpu):
    import torch

    net, lam, mu, target, t, S = _setup(gpu, seed=5)
    n = lam.shape[-1]
    pred = torch.empty((S, 36, n, n, n), dtype=torch.float32, device=gpu)
    net.forward(t["lam"], t["mu"], pred, add_eps_bar=False, extents=EXT)
    loss_dev = net.loss_backward(pred, t["target"], t["lam"], t["mu"], wm=0.3, wr=1.0)
    torch.cuda.synchronize()
    loss_gpu = float(loss_dev.item())
    loss, _, grads = _oracle(net, lam, mu, target, 0.3)
    print(f"loss gpu {loss_gpu:.6e} oracle {loss:.6e}")
    assert abs(loss_gpu - loss) <= 1e-3 * loss
    g = net.grads_dict()
    worst = 0.0
    for name, (dw, db) in grads.items():
        gw, gb = g[name]
        if name == "head":
            gw, gb, dw, db = gw[:18], gb[:18], dw[:18], db[:18]
            assert np.all(g[name][0][18:] == 0) and np.all(g[name][1][18:] == 0)
            # a constant displacement has no strain: the head bias does not
            # enter the loss (the oracle's gradient is 0; the GPU's is rounding)
            assert np.abs(db).max() <= 1e-12 * np.abs(dw).max()
            assert np.abs(gb).max() <= 1e-2 * np.abs(gw).max()
            gb, db = dw.sum(axis=(1, 2)), dw.sum(axis=(1, 2))
        ew, eb_ = _rel(gw, dw), _rel(gb, db)
        cos = float(np.dot(gw.ravel(), dw.ravel()) / (np.linalg.norm(gw) * np.linalg.norm(dw)))
        print(f"{name:6s} dW rel {ew:.2e} cos {cos:.6f}  db rel {eb_:.2e}")
        worst = max(worst, ew, eb_)
        assert ew < 3e-2 and eb_ < 3e-2 and cos > 0.999, name
    print(f"worst gradient rel L2 err {worst:.2e}")


def test_adam_reduces_loss(gpu):
    """Adam steps on one batch lower its loss toward a learnable target (zero
    fluctuation, residual term only): the optimiser and the repacking of
    updated weights work end to end."""
    import torch

    net, lam, mu, target, t, S = _setup(gpu, seed=7)
    n = lam.shape[-1]
    zero = torch.zeros_like(t["target"])
    pred = torch.empty((S, 36, n, n, n), dtype=torch.float32, device=gpu)
    losses = []
    for _ in range(12):
        net.forward(t["lam"], t["mu"], pred, add_eps_bar=False, extents=EXT)
        losses.append(float(net.loss_backward(pred, zero, t["lam"], t["mu"], wm=0.0, wr=1.0).item()))
        net.adam_step_(1e-4)
    torch.cuda.synchronize()
    print("losses", [f"{v:.4e}" for v in losses])
    assert losses[-1] < 0.5 * losses[0]


def test_invalid_arguments_are_rejected(gpu):
    import ctypes

    from jaxmaterials_amd import _lib

    lib = _lib.load()
    dims = (ctypes.c_int * 4)(1, 8, 8, 8)
    # cin not a multiple of 8: refused before any launch
    rc = lib.ls_ws_conv3d(dims, ctypes.c_void_p(16), 8, 6, ctypes.c_void_p(16), 192, None, 16, ctypes.c_void_p(16),
                          16, 0, None, 0, None, 0, None)
    assert rc == -1


def _norm_v(m):
    """Voigt norm of each row of mean stresses [R][6] (shear counted twice)."""
    m = np.asarray(m, np.float64)
    return np.sqrt(np.sum(np.array([1, 1, 1, 2, 2, 2.0]) * m ** 2, axis=-1))


def _uniform_mean_stress(lam, mu):
    """<C e_k> for the six unit load cases (rows)."""
    la, m = float(lam.double().mean()), float(mu.double().mean())
    out = np.zeros((6, 6))
    for k in range(6):
        if k < 3:
            out[k, :3] = la
            out[k, k] += 2 * m
        else:
            out[k, k] = 2 * m
    return out


@pytest.mark.parametrize("n", [64, 128])
def test_trained_warm_start_vs_uniform_start(gpu, n):
    """What the trained network (jaxmaterials_amd/warmstart/crossply64.npz,
    trained at 64^3) gains on a held-out microstructure of its family at 64^3
    and at 128^3, for conjugate gradients and Anderson depth 6: the equilibrium residual
    norm of its start relative to the uniform start's (normal load cases below
    WS_MAX_RATIO, none above 1: the solver falls back), fewer iterations in
    total than from the uniform start with CG (Anderson: at most as many,
    within 3), and the same C_eff."""
    import torch

    from jaxmaterials_amd.common import GridSpec
    from jaxmaterials_amd.data import crossply_lame
    from jaxmaterials_amd.solver.device import LSSolver, unit_load_cases
    from jaxmaterials_amd.warmstart import WarmStartNet
    from jaxmaterials_amd.warmstart.train import DEFAULT_WEIGHTS, warm_start_solve

    net = WarmStartNet.load(DEFAULT_WEIGHTS, device=gpu)
    shape, ext = (n,) * 3, (float(n),) * 3
    gs = GridSpec(*shape, *ext)
    lam, mu, _, _ = crossply_lame(shape, ext, n // 16, 6.0, 0.5, seed=7)  # training seeds start at 1 000 000
    loads = unit_load_cases()
    for cg, depth in ((True, 0), (False, 6)):
        solver = LSSolver(gs, nrhs=6, depth=depth)
        ref = solver.solve(lam, mu, loads, rtol=1e-6, maxiter=3000, cg=cg)
        start = warm_start_solve(net, solver, lam, mu, loads, rtol=1e-6, maxiter=0, check=False, cg=cg)
        # equilibrium residual norm of the kept start relative to the uniform start's:
        # R (normalised by the start's own mean stress) times the ratio of the mean stresses
        r0 = start["residual"] * _norm_v(start["sigma_mean"]) / _norm_v(_uniform_mean_stress(lam, mu))
        out = warm_start_solve(net, solver, lam, mu, loads, rtol=1e-6, maxiter=3000, cg=cg)
        C0, Cw = ref["sigma_mean"], out["sigma_mean"]
        sig = np.abs(C0) > 1e-3 * np.abs(C0).max()
        dC = float((np.abs(Cw - C0)[sig] / np.abs(C0)[sig]).max())
        print(f"{n}^3 {'CG' if cg else 'AA(6)'}: initial residual vs uniform {np.round(r0, 3).tolist()}; iterations "
              f"uniform {ref['iters'].tolist()} warm {out['iters'].tolist()}; C_eff max rel diff {dC:.2e}", flush=True)
        # the kept start is never worse than the uniform one (the solver falls back)
        assert np.all(r0 <= 1.0 + 1e-5), r0
        # the normal load cases start well below the uniform start's residual
        assert np.all(r0[:3] < WS_MAX_RATIO), r0
        if cg:
            assert out["iters"].sum() < ref["iters"].sum()
        else:
            assert out["iters"].sum() <= ref["iters"].sum() + 3
        assert out["converged"] and dC < 1e-5
        solver.close()
        torch.cuda.empty_cache()
