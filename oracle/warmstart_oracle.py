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
  * input channels (log lambda - a0) s0, (log mu - a1) s1, eps_bar (6);
    output = 6-channel strain fluctuation; loss = mean squared error.

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
    ("head", 16, 6, False),
)


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


def network_input(lam, mu, eps_bar, per_mat, norm, bf16):
    """lam, mu [nmat][X][Y][Z] float64 arrays; eps_bar [B][6]."""
    lam = torch.as_tensor(lam, dtype=torch.float64)
    mu = torch.as_tensor(mu, dtype=torch.float64)
    eb = torch.as_tensor(np.asarray(eps_bar, dtype=np.float64))
    B = eb.shape[0]
    mats = [b // per_mat for b in range(B)]
    # the GPU evaluates the normalisation in FP32
    l32 = lambda a: torch.log(a.to(torch.float32)).to(torch.float64)
    c0 = torch.stack([(l32(lam[m]) - norm[0]) * norm[1] for m in mats])
    c1 = torch.stack([(l32(mu[m]) - norm[2]) * norm[3] for m in mats])
    shape = c0.shape[1:]
    ch = [c0, c1] + [eb[:, j].view(B, 1, 1, 1).expand(B, *shape) for j in range(6)]
    x = torch.stack(ch, dim=1)
    return bf16_round(x) if bf16 else x


def forward(params, x, bf16=True):
    """params: name -> (W [cout][27][cin], b [cout]) float64 tensors (the head
    may carry extra output rows; only the first 6 are used).  x [B][8][X][Y][Z].
    Returns the fluctuation [B][6][X][Y][Z]."""
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
    w, b = w[:6], b[:6]
    if bf16:
        w = bf16_round(w)
    return conv_periodic(y1, w, b)


def loss_and_grads(np_params, lam, mu, eps_bar, per_mat, norm, target, bf16=True):
    """Mean squared error over [B][6][X][Y][Z] of forward() against target and
    its gradients.  np_params: name -> (W, b) numpy arrays.  Returns
    (loss, pred, grads) with grads: name -> (dW, db) numpy arrays of the
    shapes of np_params."""
    params = {k: (torch.tensor(np.asarray(w, np.float64), requires_grad=True),
                  torch.tensor(np.asarray(b, np.float64), requires_grad=True)) for k, (w, b) in np_params.items()}
    x = network_input(lam, mu, eps_bar, per_mat, norm, bf16)
    pred = forward(params, x, bf16)
    tgt = torch.as_tensor(np.asarray(target, np.float64))
    if bf16:
        tgt = tgt.to(torch.float32).to(torch.bfloat16).to(torch.float64)
    loss = ((pred - tgt) ** 2).mean()
    loss.backward()
    grads = {k: (w.grad.numpy(), b.grad.numpy()) for k, (w, b) in params.items()}
    return float(loss.detach()), pred.detach().numpy(), grads
