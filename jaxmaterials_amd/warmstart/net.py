"""Learned warm start of the Lippmann-Schwinger iteration: a periodic 3-D conv
encoder-decoder evaluated and trained by the HIP kernels of
csrc/conv3d.hip (BF16 MFMA, FP32 accumulation) through the C ABI of
include/ls_warmstart.h.  torch only holds device memory and the stream.

No counterpart in the reference (its iteration starts from the uniform strain,
src/jaxmaterials/solver/lippmann_schwinger.py:181-183).

Architecture: levels of 16, 32, 64, 128 channels, two 3x3x3 periodic
convolutions per level with leaky ReLU (slope 0.1), 2x average pooling down,
nearest 2x upsampling up, decoder inputs [upsampled, skip]; input
(log lambda - a0) s0, (log mu - a1) s1 (six more input channels stay zero);
output the displacement fluctuations of all six unit load cases at once, 18
channels (channel 3 k + c = component c under eps_bar = e_k): the
fluctuation is linear in eps_bar, so one evaluation per microstructure
serves its whole effective-tensor solve.  The strain is the discrete
symmetric gradient sym(D^+ u) of the rotated staggered grid
(ls_ws_symgrad), compatible by construction, 36 channels (6 k + c).
1 469 664 weights + 722 biases (the head is stored with 24 output rows, 6 of
them zero).
"""

import ctypes
import os

import numpy as np
import torch

from .. import _lib

__all__ = ["WarmStartNet", "LAYERS", "DEFAULT_NORM"]

# (name, cin, cout, leaky ReLU); decoder inputs are [upsampled, skip]
LAYERS = (
    ("enc1a", 8, 16, True), ("enc1b", 16, 16, True),
    ("enc2a", 16, 32, True), ("enc2b", 32, 32, True),
    ("enc3a", 32, 64, True), ("enc3b", 64, 64, True),
    ("bota", 64, 128, True), ("botb", 128, 128, True),
    ("dec3a", 192, 64, True), ("dec3b", 64, 64, True),
    ("dec2a", 96, 32, True), ("dec2b", 32, 32, True),
    ("dec1a", 48, 16, True), ("dec1b", 16, 16, True),
    ("head", 16, 24, False),
)
HEAD_VALID = 18  # six unit load cases x three displacement components
NOUT = 36        # six unit load cases x six strain components
ACT, DACT, PLANAR = 1, 2, 4

# log-space centre and inverse half-range of the cross-ply phase moduli
# (matrix lambda, mu = 2.593, 1.111; fibre 19.44, 29.17)
_LAM = (2.593, 19.44)
_MU = (1.111, 29.17)
DEFAULT_NORM = (
    0.5 * (np.log(_LAM[0]) + np.log(_LAM[1])), 2.0 / (np.log(_LAM[1]) - np.log(_LAM[0])),
    0.5 * (np.log(_MU[0]) + np.log(_MU[1])), 2.0 / (np.log(_MU[1]) - np.log(_MU[0])),
)


def _r32(n):
    return (n + 31) // 32 * 32


def _r16(n):
    return (n + 15) // 16 * 16


class _Layer:
    def __init__(self, name, cin, cout, act, woff, boff, toff, doff):
        self.name, self.cin, self.cout, self.act = name, cin, cout, act
        self.cs = cout  # channel stride of the output gradient
        self.kpad = _r32(27 * cin)
        self.rows_t = _r16(cout)
        self.kpad_d = _r32(27 * self.cs)
        self.rows_d = _r16(cin)
        self.woff, self.boff = woff, boff
        self.toff, self.doff = toff, doff
        self.nw = cout * 27 * cin

    @property
    def nt(self):
        return self.rows_t * self.kpad

    @property
    def nd(self):
        return self.rows_d * self.kpad_d


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"liblippmannschwinger: {what} failed ({rc})")


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


class WarmStartNet:
    """Parameters (fp32 master, bf16 packed copies) on one GPU and the
    forward/backward schedule over the C ABI."""

    def __init__(self, device=None, seed=0, norm=DEFAULT_NORM):
        self.lib = _lib.load()
        self.device = torch.device(device or "cuda:0")
        self.norm = tuple(float(v) for v in norm)
        self.layers = {}
        woff = toff = doff = 0
        for name, cin, cout, act in LAYERS:
            L = _Layer(name, cin, cout, act, woff, woff + cout * 27 * cin, toff, doff)
            woff += L.nw + cout
            toff += L.nt
            doff += L.nd
            self.layers[name] = L
        self.nparams = woff
        self.params = torch.zeros(woff, dtype=torch.float32, device=self.device)
        self.grads = torch.zeros_like(self.params)
        self.adam_m = torch.zeros_like(self.params)
        self.adam_v = torch.zeros_like(self.params)
        self.adam_step = 0
        self.wt = torch.zeros(toff, dtype=torch.int16, device=self.device)
        self.wd = torch.zeros(doff, dtype=torch.int16, device=self.device)
        self.part = torch.empty(1 << 24, dtype=torch.float32, device=self.device)
        self.part64 = torch.empty(1 << 20, dtype=torch.float64, device=self.device)
        self.loss_dev = torch.zeros(1, dtype=torch.float64, device=self.device)
        self.loss_scale = torch.zeros(1 << 12, dtype=torch.float32, device=self.device)
        self._norm_dev = torch.tensor(self.norm, dtype=torch.float32, device=self.device)
        self._acts = {}
        self.init_params(seed)

    # -- parameters -------------------------------------------------------
    def init_params(self, seed):
        """He-uniform weights (fan-in 27 cin), zero biases; head rows >= 6 zero."""
        rng = np.random.default_rng(seed)
        host = np.zeros(self.nparams, np.float32)
        for L in self.layers.values():
            bound = np.sqrt(6.0 / (27 * L.cin)) * (1.0 if L.act else 0.1)
            w = rng.uniform(-bound, bound, size=(L.cout, 27, L.cin)).astype(np.float32)
            if L.name == "head":
                w[HEAD_VALID:] = 0.0
            host[L.woff:L.woff + L.nw] = w.ravel()
        self.params.copy_(torch.from_numpy(host))
        self.pack()

    def stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def pack(self):
        st = self.stream()
        for L in self.layers.values():
            w = self.params[L.woff:]
            _check(self.lib.ls_ws_pack(_p(w), L.cout, L.cin, L.cs, _p(self.wt[L.toff:]), L.kpad, L.rows_t,
                                       _p(self.wd[L.doff:]), L.kpad_d, L.rows_d, st), "pack")

    def state_dict(self):
        """name -> (W [cout][27][cin], b [cout]) float32 numpy arrays."""
        host = self.params.cpu().numpy()
        return {L.name: (host[L.woff:L.woff + L.nw].reshape(L.cout, 27, L.cin).copy(),
                         host[L.boff:L.boff + L.cout].copy()) for L in self.layers.values()}

    def load_state_dict(self, sd):
        host = np.zeros(self.nparams, np.float32)
        for L in self.layers.values():
            w, b = sd[L.name]
            host[L.woff:L.woff + L.nw] = np.asarray(w, np.float32).reshape(-1)
            host[L.boff:L.boff + L.cout] = np.asarray(b, np.float32).reshape(-1)
        self.params.copy_(torch.from_numpy(host))
        self.pack()

    def save(self, path, **meta):
        arrays = {}
        for k, (w, b) in self.state_dict().items():
            arrays[k + ".w"] = w
            arrays[k + ".b"] = b
        arrays["norm"] = np.array(self.norm, np.float64)
        for k, v in meta.items():
            arrays["meta." + k] = np.asarray(v)
        np.savez(path, **arrays)

    @classmethod
    def load(cls, path, device=None):
        with np.load(path, allow_pickle=False) as z:
            net = cls(device=device, norm=tuple(z["norm"]))
            net.load_state_dict({L[0]: (z[L[0] + ".w"], z[L[0] + ".b"]) for L in LAYERS})
            net.meta = {k[5:]: z[k] for k in z.files if k.startswith("meta.")}
        return net

    # -- buffers ----------------------------------------------------------
    def _buf(self, key, nvox, c, dtype=torch.int16):
        t = self._acts.get(key)
        n = nvox * c
        if t is None or t.numel() < n or t.dtype != dtype:
            t = torch.empty(n, dtype=dtype, device=self.device)
            self._acts[key] = t
        return t[:n]

    def release(self):
        self._acts = {}

    # -- kernels ------------------------------------------------------------
    def _conv(self, L, dims, x, ldx, out, ldo, flags, addc=None, cvalid=0):
        dims_a = (ctypes.c_int * 4)(*dims)
        bias = self.params[L.boff:]
        _check(self.lib.ls_ws_conv3d(dims_a, _p(x), ldx, L.cin, _p(self.wt[L.toff:]), L.kpad, _p(bias), L.cout,
                                     _p(out), ldo, flags, None, 0, _p(addc), cvalid, self.stream()), L.name)

    def _dconv(self, L, dims, dz, out, ldo, tsrc=None, ldt=0):
        """Gradient with respect to the input of layer L (times act'(tsrc) if given)."""
        dims_a = (ctypes.c_int * 4)(*dims)
        flags = DACT if tsrc is not None else 0
        _check(self.lib.ls_ws_conv3d(dims_a, _p(dz), L.cs, L.cs, _p(self.wd[L.doff:]), L.kpad_d, None, L.cin,
                                     _p(out), ldo, flags, _p(tsrc), ldt, None, 0, self.stream()), L.name + " dgrad")

    def _wgrad(self, L, dims, dz, x, ldx):
        dims_a = (ctypes.c_int * 4)(*dims)
        nvox = dims[0] * dims[1] * dims[2] * dims[3]
        planar = self._buf("planar", nvox, L.cout + L.cin)
        _check(self.lib.ls_ws_conv3d_wgrad(dims_a, _p(dz), L.cs, L.cout, _p(x), ldx, L.cin, _p(planar),
                                           planar.numel(), _p(self.part), self.part.numel(),
                                           _p(self.grads[L.woff:]), _p(self.grads[L.boff:]), self.stream()),
               L.name + " wgrad")

    def _simple(self, fn, dims, *args):
        dims_a = (ctypes.c_int * 4)(*dims)
        _check(getattr(self.lib, fn)(dims_a, *args, self.stream()), fn)

    # -- forward ------------------------------------------------------------
    def forward(self, lam, mu, out, add_eps_bar=True, extents=None):
        """Predict the strains of the six unit load cases of S microstructures
        on one grid.

        lam, mu: float32 device tensors [S][X][Y][Z]; out: float32 device tensor
        [S][36][X][Y][Z] (equivalently [6 S][6][X][Y][Z], the solver's strain
        layout for the unit load cases of each microstructure), receives
        e_k + fluctuation (or the fluctuation alone if add_eps_bar is False);
        extents = (Lx, Ly, Lz) of the box (default: one unit per voxel).
        Keeps the activations for backward()."""
        X, Y, Z = lam.shape[-3:]
        if X % 8 or Y % 8 or Z % 8:
            raise ValueError("grid sizes must be multiples of 8 (three 2x poolings)")
        B = lam.reshape((-1, X, Y, Z)).shape[0]
        nper = X * Y * Z
        if out.numel() != B * NOUT * nper or out.dtype != torch.float32 or not out.is_contiguous():
            raise ValueError("out must be a contiguous float32 tensor of [S][36][X][Y][Z] elements")
        lay = self.layers
        d1, d2, d3, d4 = ((B, X >> s, Y >> s, Z >> s) for s in range(4))
        n1, n2, n3, n4 = (d[0] * d[1] * d[2] * d[3] for d in (d1, d2, d3, d4))
        st = self.stream()
        if getattr(self, "_eb0", None) is None or self._eb0.shape[0] < B:
            self._eb0 = torch.zeros((B, 6), dtype=torch.float32, device=self.device)
        eps_bar = self._eb0[:B]
        self._extents = np.asarray(extents if extents is not None else (X, Y, Z), dtype=np.float32)
        per_mat = 1
        x0 = self._buf("x0", n1, 8)
        _check(self.lib.ls_ws_input(_p(lam), _p(mu), per_mat, _p(eps_bar), B, nper, _p(self._norm_dev), _p(x0), st),
               "input")
        t1 = self._buf("t1", n1, 16)
        cat1 = self._buf("cat1", n1, 48)
        self._conv(lay["enc1a"], d1, x0, 8, t1, 16, ACT)
        self._conv(lay["enc1b"], d1, t1, 16, cat1[32:], 48, ACT)            # skip1 -> cat1[32:48]
        p1 = self._buf("p1", n2, 16)
        self._simple("ls_ws_pool", d1, _p(cat1[32:]), 48, 16, _p(p1), 16)
        t2 = self._buf("t2", n2, 32)
        cat2 = self._buf("cat2", n2, 96)
        self._conv(lay["enc2a"], d2, p1, 16, t2, 32, ACT)
        self._conv(lay["enc2b"], d2, t2, 32, cat2[64:], 96, ACT)            # skip2 -> cat2[64:96]
        p2 = self._buf("p2", n3, 32)
        self._simple("ls_ws_pool", d2, _p(cat2[64:]), 96, 32, _p(p2), 32)
        t3 = self._buf("t3", n3, 64)
        cat3 = self._buf("cat3", n3, 192)
        self._conv(lay["enc3a"], d3, p2, 32, t3, 64, ACT)
        self._conv(lay["enc3b"], d3, t3, 64, cat3[128:], 192, ACT)          # skip3 -> cat3[128:192]
        p3 = self._buf("p3", n4, 64)
        self._simple("ls_ws_pool", d3, _p(cat3[128:]), 192, 64, _p(p3), 64)
        t4 = self._buf("t4", n4, 128)
        y4 = self._buf("y4", n4, 128)
        self._conv(lay["bota"], d4, p3, 64, t4, 128, ACT)
        self._conv(lay["botb"], d4, t4, 128, y4, 128, ACT)
        self._simple("ls_ws_upsample", d3, _p(y4), 128, 128, _p(cat3), 192)  # up(y4) -> cat3[0:128]
        u3 = self._buf("u3", n3, 64)
        y3 = self._buf("y3", n3, 64)
        self._conv(lay["dec3a"], d3, cat3, 192, u3, 64, ACT)
        self._conv(lay["dec3b"], d3, u3, 64, y3, 64, ACT)
        self._simple("ls_ws_upsample", d2, _p(y3), 64, 64, _p(cat2), 96)
        u2 = self._buf("u2", n2, 32)
        y2 = self._buf("y2", n2, 32)
        self._conv(lay["dec2a"], d2, cat2, 96, u2, 32, ACT)
        self._conv(lay["dec2b"], d2, u2, 32, y2, 32, ACT)
        self._simple("ls_ws_upsample", d1, _p(y2), 32, 32, _p(cat1), 48)
        u1 = self._buf("u1", n1, 16)
        y1 = self._buf("y1", n1, 16)
        self._conv(lay["dec1a"], d1, cat1, 48, u1, 16, ACT)
        self._conv(lay["dec1b"], d1, u1, 16, y1, 16, ACT)
        u = self._buf("u_head", n1, HEAD_VALID, dtype=torch.float32)
        self._conv(lay["head"], d1, y1, 16, u, 0, PLANAR, cvalid=HEAD_VALID)
        dims_a = (ctypes.c_int * 4)(B, X, Y, Z)
        _check(self.lib.ls_ws_symgrad(dims_a, self._extents.ctypes.data_as(ctypes.c_void_p), _p(u),
                                      1 if add_eps_bar else 0, _p(out), st), "symgrad")
        self._dims = (d1, d2, d3, d4)
        return out

    # -- backward -----------------------------------------------------------
    def loss_backward(self, pred, target, lam, mu, wm=1.0, wr=0.0, backward=True):
        """Loss of pred (float32 [S][36][X][Y][Z], the strain fluctuation from
        forward(..., add_eps_bar=False)) against target (int16-stored bf16, same
        shape) on the solver's own terms (ls_ws_residual_loss: wm times the
        relative strain energy of the error, the norm conjugate gradients
        reduce, plus wr times the squared equilibrium residual of the predicted
        strain relative to the uniform start's), for the Lame fields
        lam, mu [S][X][Y][Z] on the box of the last forward(); fills self.grads
        (with backward).  Returns the loss (device tensor)."""
        d1, d2, d3, d4 = self._dims
        n1, n2, n3, n4 = (d[0] * d[1] * d[2] * d[3] for d in (d1, d2, d3, d4))
        S, nper = d1[0], d1[1] * d1[2] * d1[3]
        st = self.stream()
        lay = self.layers
        A = self._acts
        hc = lay["head"].cout
        ge = self._buf("g_eps", n1, 40)
        scratch = self._buf("rl_scratch", S * nper, 54, dtype=torch.float32)
        nb = (nper + 255) // 256
        if 24 * S * nb > self.part64.numel():
            self.part64 = torch.empty(24 * S * nb, dtype=torch.float64, device=self.device)
        if 12 * S > self.loss_scale.numel():
            self.loss_scale = torch.zeros(12 * S, dtype=torch.float32, device=self.device)
        dims_a = (ctypes.c_int * 4)(S, d1[1], d1[2], d1[3])
        ext = self._extents.ctypes.data_as(ctypes.c_void_p)
        _check(self.lib.ls_ws_residual_loss(dims_a, ext, _p(pred), _p(target), _p(lam), _p(mu), float(wm), float(wr),
                                            _p(ge), 40,
                                            _p(scratch), scratch.numel(), _p(self.part64), self.part64.numel(),
                                            _p(self.loss_scale), _p(self.loss_dev), st), "residual loss")
        if not backward:
            return self.loss_dev
        gh = self._buf("g_head", n1, hc)
        _check(self.lib.ls_ws_symgrad_backward(dims_a, ext, _p(ge), 40, _p(gh), hc, st), "symgrad backward")
        self.grads.zero_()
        # decoder level 1 and head
        self._wgrad(lay["head"], d1, gh, A["y1"], 16)
        dy1 = self._buf("dy1", n1, 16)
        self._dconv(lay["head"], d1, gh, dy1, 16, A["y1"], 16)
        self._wgrad(lay["dec1b"], d1, dy1, A["u1"], 16)
        du1 = self._buf("du1", n1, 16)
        self._dconv(lay["dec1b"], d1, dy1, du1, 16, A["u1"], 16)
        self._wgrad(lay["dec1a"], d1, du1, A["cat1"], 48)
        dcat1 = self._buf("dcat1", n1, 48)
        self._dconv(lay["dec1a"], d1, du1, dcat1, 48, A["cat1"], 48)
        # decoder level 2
        dy2 = self._buf("dy2", n2, 32)
        self._simple("ls_ws_upsample_backward", d1, _p(dcat1), 48, 32, _p(dy2), 32)
        self._wgrad(lay["dec2b"], d2, dy2, A["u2"], 32)
        du2 = self._buf("du2", n2, 32)
        self._dconv(lay["dec2b"], d2, dy2, du2, 32, A["u2"], 32)
        self._wgrad(lay["dec2a"], d2, du2, A["cat2"], 96)
        dcat2 = self._buf("dcat2", n2, 96)
        self._dconv(lay["dec2a"], d2, du2, dcat2, 96, A["cat2"], 96)
        # decoder level 3
        dy3 = self._buf("dy3", n3, 64)
        self._simple("ls_ws_upsample_backward", d2, _p(dcat2), 96, 64, _p(dy3), 64)
        self._wgrad(lay["dec3b"], d3, dy3, A["u3"], 64)
        du3 = self._buf("du3", n3, 64)
        self._dconv(lay["dec3b"], d3, dy3, du3, 64, A["u3"], 64)
        self._wgrad(lay["dec3a"], d3, du3, A["cat3"], 192)
        dcat3 = self._buf("dcat3", n3, 192)
        self._dconv(lay["dec3a"], d3, du3, dcat3, 192, A["cat3"], 192)
        # bottleneck
        dy4 = self._buf("dy4", n4, 128)
        self._simple("ls_ws_upsample_backward", d3, _p(dcat3), 192, 128, _p(dy4), 128)
        self._wgrad(lay["botb"], d4, dy4, A["t4"], 128)
        dt4 = self._buf("dt4", n4, 128)
        self._dconv(lay["botb"], d4, dy4, dt4, 128, A["t4"], 128)
        self._wgrad(lay["bota"], d4, dt4, A["p3"], 64)
        dp3 = self._buf("dp3", n4, 64)
        self._dconv(lay["bota"], d4, dt4, dp3, 64)
        # encoder level 3
        ds3 = self._buf("ds3", n3, 64)
        self._simple("ls_ws_pool_backward", d3, _p(dp3), 64, 64, _p(A["cat3"][128:]), 192, _p(dcat3[128:]), 192,
                     _p(ds3), 64)
        self._wgrad(lay["enc3b"], d3, ds3, A["t3"], 64)
        dt3 = self._buf("dt3", n3, 64)
        self._dconv(lay["enc3b"], d3, ds3, dt3, 64, A["t3"], 64)
        self._wgrad(lay["enc3a"], d3, dt3, A["p2"], 32)
        dp2 = self._buf("dp2", n3, 32)
        self._dconv(lay["enc3a"], d3, dt3, dp2, 32)
        # encoder level 2
        ds2 = self._buf("ds2", n2, 32)
        self._simple("ls_ws_pool_backward", d2, _p(dp2), 32, 32, _p(A["cat2"][64:]), 96, _p(dcat2[64:]), 96,
                     _p(ds2), 32)
        self._wgrad(lay["enc2b"], d2, ds2, A["t2"], 32)
        dt2 = self._buf("dt2", n2, 32)
        self._dconv(lay["enc2b"], d2, ds2, dt2, 32, A["t2"], 32)
        self._wgrad(lay["enc2a"], d2, dt2, A["p1"], 16)
        dp1 = self._buf("dp1", n2, 16)
        self._dconv(lay["enc2a"], d2, dt2, dp1, 16)
        # encoder level 1
        ds1 = self._buf("ds1", n1, 16)
        self._simple("ls_ws_pool_backward", d1, _p(dp1), 16, 16, _p(A["cat1"][32:]), 48, _p(dcat1[32:]), 48,
                     _p(ds1), 16)
        self._wgrad(lay["enc1b"], d1, ds1, A["t1"], 16)
        dt1 = self._buf("dt1", n1, 16)
        self._dconv(lay["enc1b"], d1, ds1, dt1, 16, A["t1"], 16)
        self._wgrad(lay["enc1a"], d1, dt1, A["x0"], 8)
        return self.loss_dev

    def adam_step_(self, lr, b1=0.9, b2=0.999, eps=1e-8):
        self.adam_step += 1
        _check(self.lib.ls_ws_adam(_p(self.params), _p(self.grads), _p(self.adam_m), _p(self.adam_v), self.nparams,
                                   lr, b1, b2, eps, self.adam_step, self.stream()), "adam")
        self.pack()

    def grads_dict(self):
        host = self.grads.cpu().numpy()
        return {L.name: (host[L.woff:L.woff + L.nw].reshape(L.cout, 27, L.cin).copy(),
                         host[L.boff:L.boff + L.cout].copy()) for L in self.layers.values()}

    def make_target(self, eps, eps_bar, out=None):
        """bf16 (int16 storage) target eps - eps_bar from solver strains [B][6][...]."""
        B = eps.shape[0]
        nper = eps[0, 0].numel()
        if out is None:
            out = torch.empty(eps.shape, dtype=torch.int16, device=self.device)
        _check(self.lib.ls_ws_target(_p(eps), _p(eps_bar), B, nper, _p(out), self.stream()), "target")
        return out
