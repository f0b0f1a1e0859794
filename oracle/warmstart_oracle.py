"""CPU FP64 oracle of the warm-start network (test infrastructure only).

The reference has no warm start (its iteration always starts from the
uniform strain, src/jaxmaterials/solver/lippmann_schwinger.py:181-183); this
is the specification of the network that the HIP kernels of
jaxmaterials_amd/csrc/conv3d.hip evaluate and train, restated with torch's
CPU float64 operators so that autograd gives the reference gradients:

  * 3x3x3 periodic convolutions (circular padding, cross-correlation: the
    output at x reads the input at x + (dx, dy, dz), dx in {-1, 0, 1}), weight
    W[cout][27][cin] with tap = 9 (dx+1) + 3 (dy+1) + (dz+1), plus bias;
  * leaky ReLU (slope 0.1) after every convolution but the last;
  * 2x average pooling, nearest 2x upsampling, decoder inputs concatenate
    [upsampled, skip] along channels;
  * levels of 16, 32, 64, 128 channels, two convolutions per level;
  * input channels (log lambda - a0) s0, (log mu - a1) s1 and six zero
    channels; output = the displacement fluctuations of the six unit load
    cases, 18 channels (3 k + c: component c under eps_bar = e_k), and the
    strain is their discrete symmetric gradient sym(D^+ u) (symgrad), 36
    channels (6 k + c);
  * loss (residual_loss) = wm times the relative strain energy of the error
    plus wr times the squared equilibrium residual of the predicted strain
    relative to the uniform start's, with the backward divergence of the
    rotated staggered grid (src/jaxmaterials/solver/derivatives.py:8-96).

With ``bf16=True`` the oracle rounds to bfloat16 where the GPU stores bf16
(input, weights, every activation, pooled tensors, the target), with
straight-through gradients, so forward values are comparable to the GPU to
FP32-accumulation accuracy; the backward pass stays FP64.

Pinned in tests/test_oracle_warmstart.py: the periodic convolution against an
explicit loop, pooling/upsampling against loops, the parameter count, and
autograd against central finite differences.
"""

import numpy as np
import torch
import torch.nn.functional as F

SLOPE = 0.1

# (name, cin, cout, leaky ReLU)
LAYERS = (
    ("enc1a", 8, 16, True), ("enc1b", 16, 16, True),
    ("enc2a", 16, 32, True), ("enc2b", 32, 32, True),
    ("enc3a", 32, 64, True), ("enc3b", 64, 64, True),
    ("bota", 64, 128, True), ("botb", 128, 128, True),
    ("dec3a", 192, 64, True), ("dec3b", 64, 64, True),
    ("dec2a", 96, 32, True), ("dec2b", 32, 32, True),
    ("dec1a", 48, 16, True), ("dec1b", 16, 16, True),
    ("head", 16, 18, False),
)
HEAD = 18


def param_count():
    return sum(27 * ci * co + co for _, ci, co, _ in LAYERS)


def bf16_round(x):
    """Round to bfloat16 (via float32, as the GPU rounds its FP32 results),
    straight-through in the backward pass."""
    r = x.detach().to(torch.float32).to(torch.bfloat16).to(x.dtype)
    return x + (r - x.detach())


def conv_periodic(x, w27, b):
    """x [B][cin][X][Y][Z]; w27 [cout][27][cin]; periodic 3x3x3 correlation."""
    co, _, ci = w27.shape
    w = w27.reshape(co, 3, 3, 3, ci).permute(0, 4, 1, 2, 3)
    xp = F.pad(x, (1, 1, 1, 1, 1, 1), mode="circular")
    return F.conv3d(xp, w, b)


def conv_periodic_loop(x, w27, b):
    """Explicit-loop restatement of conv_periodic (numpy, for tiny inputs)."""
    B, ci, X, Y, Z = x.shape
    co = w27.shape[0]
    out = np.zeros((B, co, X, Y, Z))
    for tap in range(27):
        dx, dy, dz = tap // 9 - 1, (tap // 3) % 3 - 1, tap % 3 - 1
        shifted = np.roll(x, shift=(-dx, -dy, -dz), axis=(2, 3, 4))  # shifted[..., i] = x[..., i + d]
        out += np.einsum("oc,bcxyz->boxyz", w27[:, tap, :], shifted)
    return out + b[None, :, None, None, None]


def pool(x):
    return F.avg_pool3d(x, 2)


def upsample(x):
    return x.repeat_interleave(2, dim=2).repeat_interleave(2, dim=3).repeat_interleave(2, dim=4)


def leaky(x):
    return torch.where(x > 0, x, SLOPE * x)


def network_input(lam, mu, norm, bf16):
    """lam, mu [S][X][Y][Z] float64 arrays (one instance per microstructure)."""
    lam = torch.as_tensor(lam, dtype=torch.float64)
    mu = torch.as_tensor(mu, dtype=torch.float64)
    # the GPU evaluates the normalisation in FP32
    l32 = lambda a: torch.log(a.to(torch.float32)).to(torch.float64)
    c0 = (l32(lam) - norm[0]) * norm[1]
    c1 = (l32(mu) - norm[2]) * norm[3]
    zero = torch.zeros_like(c0)
    x = torch.stack([c0, c1] + [zero] * 6, dim=1)
    return bf16_round(x) if bf16 else x


def forward(params, x, bf16=True):
    """params: name -> (W [cout][27][cin], b [cout]) float64 tensors (the head
    may carry extra output rows; only the first 18 are used).  x [S][8][X][Y][Z].
    Returns the displacement fluctuations [S][18][X][Y][Z]."""
    rnd = bf16_round if bf16 else (lambda t: t)

    def conv(name, t, act=True):
        w, b = params[name]
        if bf16:
            w = bf16_round(w)
        y = conv_periodic(t, w, b)
        return rnd(leaky(y)) if act else y

    s1 = conv("enc1b", conv("enc1a", x))
    s2 = conv("enc2b", conv("enc2a", rnd(pool(s1))))
    s3 = conv("enc3b", conv("enc3a", rnd(pool(s2))))
    y4 = conv("botb", conv("bota", rnd(pool(s3))))
    y3 = conv("dec3b", conv("dec3a", torch.cat([upsample(y4), s3], dim=1)))
    y2 = conv("dec2b", conv("dec2a", torch.cat([upsample(y3), s2], dim=1)))
    y1 = conv("dec1b", conv("dec1a", torch.cat([upsample(y2), s1], dim=1)))
    w, b = params["head"]
    w, b = w[:HEAD], b[:HEAD]
    if bf16:
        w = bf16_round(w)
    return conv_periodic(y1, w, b)


def divergence(sig, extents):
    """Backward divergence D^- . sigma of Voigt stresses [..., 6, X, Y, Z]
    (xx yy zz xy xz yz) on the rotated staggered grid: D^-_d f = (n_d / (4 L_d))
    (1 + S_e)(1 + S_f)(1 - S_d) f, S the periodic shift i -> i - 1
    (derivatives.py:8-96).  Returns [..., 3, X, Y, Z]."""
    dims = (-3, -2, -1)
    n = sig.shape[-3:]

    def dm(f, d):
        g = f
        for e in dims:
            if e != dims[d]:
                g = g + torch.roll(g, 1, e)
        return (0.25 * n[d] / extents[d]) * (g - torch.roll(g, 1, dims[d]))

    s = [sig[..., c, :, :, :] for c in range(6)]
    return torch.stack([dm(s[0], 0) + dm(s[3], 1) + dm(s[4], 2),
                        dm(s[3], 0) + dm(s[1], 1) + dm(s[5], 2),
                        dm(s[4], 0) + dm(s[5], 1) + dm(s[2], 2)], dim=-4)


def symgrad(u, extents):
    """sym(D^+ u) per load case: u [S][18][X][Y][Z] (3 k + c) -> strains
    [S][36][X][Y][Z] (6 k + c), D^+_d f = (n_d / (4 L_d)) (1 + S'_e)(1 + S'_f)(S'_d - 1) f
    with S' the periodic shift i -> i + 1 (D^+ = -(D^-)^T)."""
    dims = (-3, -2, -1)
    n = u.shape[-3:]

    def dp(f, d):
        g = f
        for e in dims:
            if e != dims[d]:
                g = g + torch.roll(g, -1, e)
        return (0.25 * n[d] / extents[d]) * (torch.roll(g, -1, dims[d]) - g)

    out = []
    for k in range(6):
        ux, uy, uz = (u[:, 3 * k + c] for c in range(3))
        out += [dp(ux, 0), dp(uy, 1), dp(uz, 2), 0.5 * (dp(ux, 1) + dp(uy, 0)), 0.5 * (dp(ux, 2) + dp(uz, 0)),
                0.5 * (dp(uy, 2) + dp(uz, 1))]
    return torch.stack(out, dim=1)


def hooke(lam, mu, e):
    """Voigt stress of strains e [..., 6, X, Y, Z] with fields lam, mu [X, Y, Z] broadcast."""
    tr = e[..., 0:1, :, :, :] + e[..., 1:2, :, :, :] + e[..., 2:3, :, :, :]
    s = 2.0 * mu.unsqueeze(-4) * e
    return torch.cat([s[..., :3, :, :, :] + lam.unsqueeze(-4) * tr, s[..., 3:, :, :, :]], dim=-4)


def residual_loss(pred, target, lam, mu, extents, wm, wr=1.0):
    """sum_{s,k} [wm <e, C e>_w / <t, C t>_w + wr |D^- . C e|^2 / |D^- . C e_k|^2] / (6 S),
    e = pred - target; pred, target [S][36][X][Y][Z], lam, mu [S][X][Y][Z]
    (torch float64).  The first term is the relative strain energy of the
    error; with an equilibrium target the second is the squared equilibrium
    residual of eps = e_k + pred relative to that of the uniform start e_k."""
    S = pred.shape[0]
    shape = pred.shape[-3:]
    w = torch.tensor([1.0, 1.0, 1.0, 2.0, 2.0, 2.0], dtype=pred.dtype).view(1, 6, 1, 1, 1)
    tot = 0.0
    for k in range(6):
        e = pred[:, 6 * k:6 * k + 6] - target[:, 6 * k:6 * k + 6]
        t = target[:, 6 * k:6 * k + 6]
        ek = torch.zeros((S, 6) + tuple(shape), dtype=pred.dtype)
        ek[:, k] = 1.0
        se, st = hooke(lam, mu, e), hooke(lam, mu, t)
        q = divergence(se, extents)
        q0 = divergence(hooke(lam, mu, ek), extents)
        tot = tot + (wm * (w * e * se).sum(dim=(1, 2, 3, 4)) / (w * t * st).sum(dim=(1, 2, 3, 4))
                     + wr * (q ** 2).sum(dim=(1, 2, 3, 4)) / (q0 ** 2).sum(dim=(1, 2, 3, 4))).sum()
    return tot / (6 * S)


def loss_and_grads(np_params, lam, mu, norm, target, extents, wm=1.0, wr=1.0, bf16=True):
    """residual_loss of forward() against target and its gradients.
    np_params: name -> (W, b) numpy arrays; lam, mu [S][X][Y][Z]; target
    [S][36][X][Y][Z].  Returns (loss, pred strain fluctuation, grads) with
    grads: name -> (dW, db) numpy arrays of the shapes of np_params."""
    params = {k: (torch.tensor(np.asarray(w, np.float64), requires_grad=True),
                  torch.tensor(np.asarray(b, np.float64), requires_grad=True)) for k, (w, b) in np_params.items()}
    x = network_input(lam, mu, norm, bf16)
    pred = symgrad(forward(params, x, bf16), extents)
    tgt = torch.as_tensor(np.asarray(target, np.float64))
    if bf16:
        tgt = tgt.to(torch.float32).to(torch.bfloat16).to(torch.float64)
    loss = residual_loss(pred, tgt, torch.as_tensor(np.asarray(lam, np.float64)),
                         torch.as_tensor(np.asarray(mu, np.float64)), extents, wm, wr)
    loss.backward()
    grads = {k: (w.grad.numpy(), b.grad.numpy()) for k, (w, b) in params.items()}
    return float(loss.detach()), pred.detach().numpy(), grads
