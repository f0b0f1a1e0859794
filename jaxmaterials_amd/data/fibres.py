"""The reference's layered-fibre dataset generator on the MI355X.

Same classes, arguments and random streams as the reference:

* ``FibreRadiusDistribution`` and ``FibreDistribution2d``
  (src/jaxmaterials/data/distributions_fibres.py:26-295); the fibre movement
  runs in liblippmannschwinger.so (``move_fibres``, the reference's
  libfibres.cc ABI, bit-identical results) instead of a g++ build at run time;
* ``LayeredFibresDatasetGenerator`` (src/jaxmaterials/data/data.py:64-331):
  ``generate_fibre_positions`` is the reference's placement (data.py:188-265),
  ``material_properties`` rasterises on the GPU (``ls_rasterise_layered_fibres``,
  bit-identical to data.py:267-314 in float32 and float64) and ``generate``
  solves the six load cases of many samples per batched solver call.

Differences, each deliberate:

* ``FibreDistribution2d(fast_code=False)`` raises: the reference's Python
  fallback is a different random stream (distributions_fibres.py:212-275),
  and this package has no host fallback of the movement;
* ``generate`` returns ``(lame_parameters, sigma_bar)``; the reference's ends
  with ``return np.asarray(data)`` on an undefined name (data.py:187) and
  raises NameError after filling the arrays;
* ``save_hdf5`` and ``LayeredFibresDataset`` need h5py, which this
  environment lacks; they are not provided.
"""

import ctypes
import itertools
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import torch
from scipy.stats import norm

from .. import _lib
from ..solver.device import LSSolver

__all__ = ["FibreRadiusDistribution", "FibreDistribution2d", "LayeredFibresDatasetGenerator",
           "rasterise_layered_fibres"]

# random-state buffer of the movement (the reference passes 10000 bytes,
# distributions_fibres.py:161-163; std::mt19937's text is below 7000)
_STATE_BYTES = 10000
# seed of the movement's Mersenne Twister (distributions_fibres.py:160)
MOVEMENT_SEED = 735283


def _dptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class FibreRadiusDistribution:
    """Distribution of fibre radii: a normal law clipped to [r_min, r_max],
    sampled by inverse transform on a 1000-point tabulated CDF, or constant
    (distributions_fibres.py:26-74)."""

    def __init__(self, r_avg, r_min, r_max, sigma, gaussian, seed=856219):
        """:arg r_avg: average fibre radius
        :arg r_min: minimum fibre radius
        :arg r_max: maximum fibre radius
        :arg sigma: standard deviation of fibre radius
        :arg gaussian: draw from the clipped normal law (else all r_avg)
        :arg seed: seed of the distribution's own numpy Generator
        """
        self.r_avg = r_avg
        self.r_min = r_min
        self.r_max = r_max
        self.sigma = sigma
        self.gaussian = gaussian
        self._rng = np.random.default_rng(seed=seed)
        nodes = np.linspace(self.r_min, self.r_max, num=1000)
        cdf = np.cumsum(norm.pdf(nodes, loc=self.r_avg, scale=self.sigma))
        self._cdf = cdf / cdf[-1]
        self._x_cdf = np.linspace(0, 1, 1000)

    def draw(self, n_samples):
        """Radii of n_samples fibres (float64 array)."""
        if not self.gaussian:
            return np.ones(shape=(n_samples,)) * self.r_avg
        u = self._rng.uniform(low=0, high=1, size=n_samples)
        return np.interp(u, self._cdf, self._x_cdf) * (self.r_max - self.r_min) + self.r_min


class FibreDistribution2d:
    """Fibre cross-sections in a periodic 2-D box by Yang Chen's random
    movement from a regular grid (distributions_fibres.py:75-295).  Iterating
    yields (locations [n, 2], radii [n]) with a fresh radius draw and movement
    each time; the movement's random state carries over between draws."""

    def __init__(self, domain_size, volume_fraction=0.55,
                 r_fibre_dist=FibreRadiusDistribution(r_avg=7.5e-3, r_min=5.0e-3, r_max=10.0e-3, sigma=0.5e-3,
                                                      gaussian=True),
                 seed=141517, fast_code=True):
        """:arg domain_size: [L_x, L_y] of the box
        :arg volume_fraction: fibre volume fraction of the starting grid
        :arg r_fibre_dist: FibreRadiusDistribution (the default instance is
            shared between calls, as in the reference)
        :arg seed: seed of a numpy Generator kept for interface parity (the
            movement uses its own Mersenne Twister)
        :arg fast_code: must be True (the native movement)
        """
        if not fast_code:
            raise NotImplementedError("FibreDistribution2d: only the native fibre movement (fast_code=True) "
                                      "is provided")
        self.domain_size = np.asarray(domain_size, dtype=np.float64)
        self._volume_fraction = volume_fraction
        self._r_fibre_dist = r_fibre_dist
        self._rng = np.random.default_rng(seed=seed)
        self._fast_code = True
        n_dir = np.ceil(self.domain_size / r_fibre_dist.r_avg * np.sqrt(volume_fraction / np.pi)).astype(np.int64)
        d = 2 * r_fibre_dist.r_avg
        axes = [self.domain_size[k] * np.arange(0, (n_dir[k] - 0.5) * d, d) / (d * n_dir[k]) for k in (0, 1)]
        self._initial_fibre_locations = np.asarray(list(itertools.product(*axes))).reshape(
            n_dir[0] * n_dir[1], 2)
        self._lib = _lib.load()
        self._rng_state_length = ctypes.c_uint()
        self._rng_state = ctypes.create_string_buffer(_STATE_BYTES)
        self._lib.initialise_rng(MOVEMENT_SEED, ctypes.byref(self._rng_state_length), self._rng_state)

    def draw_radii(self):
        """Radii of the next draw (one per starting-grid fibre)."""
        return np.ascontiguousarray(self._r_fibre_dist.draw(self._initial_fibre_locations.shape[0]),
                                    dtype=np.float64)

    def move(self, radii, num_repeats=30, it_max_ovlap=5000, eps_fibres=3.0e-4):
        """Fibre locations after the movement from the starting grid, carrying
        this instance's random state forward (distributions_fibres.py:190-211)."""
        loc = np.array(self._initial_fibre_locations, dtype=np.float64)
        self._lib.move_fibres(radii.shape[0], _dptr(self.domain_size), _dptr(radii), _dptr(loc), num_repeats,
                              eps_fibres, it_max_ovlap, ctypes.byref(self._rng_state_length), self._rng_state)
        return loc

    def __iter__(self):
        while True:
            r = self.draw_radii()
            yield (self.move(r), r)


def rasterise_layered_fibres(grid_spec, d_void, fibre_positions, fibre_radii, fibre_orientations, mu3, lambda3,
                             dtype=np.float32, out=None, device="cuda"):
    """Lame fields [mu, lambda] of layered fibres on the GPU (data.py:267-314).

    :arg fibre_positions, fibre_radii, fibre_orientations: per-layer lists as
        generate_fibre_positions returns them
    :arg mu3, lambda3: (void, fibre, matrix) values
    :arg dtype: np.float32 or np.float64
    :arg out: optional GPU tensor [2, nx, ny, nz] of that dtype
    Returns the GPU tensor [2, nx, ny, nz] (channel 0 mu, channel 1 lambda).
    """
    lib = _lib.load()
    shape = (grid_spec.nx, grid_spec.ny, grid_spec.nz)
    tdt = {np.dtype(np.float32): torch.float32, np.dtype(np.float64): torch.float64}.get(np.dtype(dtype))
    if tdt is None:
        raise TypeError("rasterise_layered_fibres: dtype must be float32 or float64")
    if out is None:
        out = torch.empty((2,) + shape, dtype=tdt, device=device)
    if not out.is_cuda or out.dtype != tdt or tuple(out.shape) != (2,) + shape or not out.is_contiguous():
        raise ValueError(f"out must be a contiguous GPU tensor {(2,) + shape} of {tdt}")
    rows = []
    for pos, rad, o in zip(fibre_positions, fibre_radii, fibre_orientations):
        pos = np.asarray(pos, dtype=np.float64).reshape(-1, 2)
        rad = np.asarray(rad, dtype=np.float64).reshape(-1)
        rows.append(np.column_stack([pos, rad, np.full(rad.shape, 1.0 if o else 0.0)]))
    fib = np.ascontiguousarray(np.concatenate(rows) if rows else np.zeros((0, 4)), dtype=np.float64)
    vox = np.array(shape, dtype=np.int32)
    ext = np.array([grid_spec.Lx, grid_spec.Ly, grid_spec.Lz], dtype=np.float64)
    m3 = np.array(mu3, dtype=np.float64)
    l3 = np.array(lambda3, dtype=np.float64)
    stream = torch.cuda.current_stream(out.device)
    rc = lib.ls_rasterise_layered_fibres(_dptr(vox), _dptr(ext), float(d_void), fib.shape[0], _dptr(fib),
                                         _dptr(m3), _dptr(l3), out.element_size(), ctypes.c_void_p(out.data_ptr()),
                                         ctypes.c_void_p(stream.cuda_stream))
    if rc != 0:
        raise RuntimeError("ls_rasterise_layered_fibres failed (see stderr)")
    return out


class LayeredFibresDatasetGenerator:
    """Samples of alternating fibre layers between two void slabs and their
    effective 6x6 tensors (data.py:64-186).  Constructor arguments as in the
    reference; ``device`` names the GPU."""

    def __init__(self, grid_spec, nlayers, d_void, n_samples, radius_distribution, volume_fraction=0.3,
                 mu_fibre=1.0, mu_material=0.5, mu_void=0.1, lambda_fibre=0.1, lambda_material=0.1,
                 lambda_void=0.01, rng=None, dtype=np.float64, verbose=False, device="cuda"):
        self.grid_spec = grid_spec
        self.nlayers = nlayers
        self.d_void = d_void
        self.n_samples = n_samples
        self.radius_distribution = radius_distribution
        self.volume_fraction = volume_fraction
        self.mu_fibre = mu_fibre
        self.mu_material = mu_material
        self.mu_void = mu_void
        self.lambda_fibre = lambda_fibre
        self.lambda_material = lambda_material
        self.lambda_void = lambda_void
        self.rng = np.random.default_rng(seed=64217) if rng is None else rng
        self.dtype = dtype
        self.verbose = verbose
        self.device = device

    def generate_fibre_positions(self):
        """(layer boundaries [nlayers + 1], positions per layer [n_j, 2],
        radii per layer [n_j], orientations per layer), data.py:188-265:
        random layer depths of at least 3 r_max between the void slabs,
        alternating orientations from a random start, fibres moved in each
        layer's box [L_c, depth - 2 r_max], copies of the fibres within r_max
        of either in-plane edge appended one period over, then lifted by the
        layer's lower boundary + r_max."""
        bounds, plan = self._draw_layers()
        for layer in plan:
            layer["moved"] = layer["dist"].move(layer["radii"])
        return self._finish_layers(bounds, plan)

    def _draw_layers(self):
        """The random draws of one sample in the reference's order (layer
        boundaries, orientation offset, then per layer the unused seed draw
        and the radii); the movements are left to the caller.  The movement of
        every layer starts from a freshly seeded Mersenne Twister, so the
        layers (and samples) can be moved in any order or in parallel."""
        gs = self.grid_spec
        r_max = self.radius_distribution.r_max
        depths = np.asarray([0.5 * r_max])
        while np.any(depths < 3 * r_max):
            inner = self.rng.uniform(size=self.nlayers - 1, low=self.d_void, high=gs.Lz - self.d_void)
            bounds = np.asarray(sorted([self.d_void, gs.Lz - self.d_void] + [float(z) for z in inner]))
            depths = bounds[1:] - bounds[:-1]
        offset = self.rng.integers(low=0, high=2, size=1)[0]
        extents = [gs.Lx, gs.Ly, gs.Lz]
        plan = []
        for k in range(self.nlayers):
            o = (k + offset) % 2
            seed = self.rng.integers(low=0, high=2, size=1)[0]
            dist = FibreDistribution2d(domain_size=[extents[o], depths[k] - 2 * r_max],
                                       volume_fraction=self.volume_fraction,
                                       r_fibre_dist=self.radius_distribution, seed=seed, fast_code=True)
            plan.append({"orientation": o, "extent": extents[o], "dist": dist,
                         "radii": dist.draw_radii()})
        return bounds, plan

    def _finish_layers(self, bounds, plan):
        """Boundary copies and vertical lift of the moved layers (data.py:239-265)."""
        r_max = self.radius_distribution.r_max
        orientations, positions, radii = [], [], []
        for k, layer in enumerate(plan):
            o, L = layer["orientation"], layer["extent"]
            X, R = layer["moved"], layer["radii"]
            orientations.append(o)
            lo = X[..., 0] <= r_max
            hi = X[..., 0] >= L - r_max
            X = np.concatenate([X, X[lo] + np.asarray([L, 0]), X[hi] - np.asarray([L, 0])], axis=0)
            R = np.concatenate([R, R[lo], R[hi]], axis=0)
            X += np.asarray([0, bounds[k] + r_max])
            positions.append(X)
            radii.append(R)
        return bounds, positions, radii, orientations

    def place_samples(self, n, threads=8):
        """generate_fibre_positions() of the next n samples, the random draws in
        sample order on this thread and the fibre movements on ``threads``
        host threads (the library call releases the GIL).  Same results as n
        sequential calls."""
        drawn = [self._draw_layers() for _ in range(n)]
        layers = [layer for _, plan in drawn for layer in plan]

        def move(layer):
            layer["moved"] = layer["dist"].move(layer["radii"])

        if threads > 1 and len(layers) > 1:
            with ThreadPoolExecutor(max_workers=threads) as pool:
                list(pool.map(move, layers))
        else:
            for layer in layers:
                move(layer)
        return [self._finish_layers(bounds, plan) for bounds, plan in drawn]

    def _triples(self):
        return ((self.mu_void, self.mu_fibre, self.mu_material),
                (self.lambda_void, self.lambda_fibre, self.lambda_material))

    def material_properties_device(self, fibre_positions, fibre_radii, fibre_orientations, out=None):
        """[mu, lambda] fields as a GPU tensor [2, nx, ny, nz] of self.dtype."""
        mu3, lam3 = self._triples()
        return rasterise_layered_fibres(self.grid_spec, self.d_void, fibre_positions, fibre_radii,
                                        fibre_orientations, mu3, lam3, dtype=self.dtype, out=out,
                                        device=self.device)

    def material_properties(self, fibre_positions, fibre_radii, fibre_orientations):
        """[mu, lambda] fields as a numpy array [2, nx, ny, nz] (data.py:267-314)."""
        return self.material_properties_device(fibre_positions, fibre_radii, fibre_orientations).cpu().numpy()

    def generate(self, rtol=1.0e-4, atol=1.0e-4, maxiter=100, depth=0, swap_lame=True, max_batch=None,
                 budget_bytes=48 << 30, placement_threads=8):
        """Fill (n_samples, 2, nx, ny, nz) Lame fields and (n_samples, 6, 6)
        effective tensors (row k = <sigma> under eps_bar = e_k), data.py:148-186.

        The reference solves each load case by one call of its loader with
        rtol = atol = 1e-4, maxiter = 100, and that loader passes mu in the
        solver's lambda slot (src/jaxmaterials/solver/lippmann_schwinger.py:310-312);
        ``swap_lame=True`` reproduces that, so the tensors match the
        reference's datasets.  Here all load cases of up to ``max_batch``
        samples go into one batched solve (``depth`` = Anderson depth; 0 is the
        reference's plain iteration).  Raises RuntimeError if a load case does
        not converge within maxiter, like the reference's loader.
        Returns (lame_parameters, sigma_bar) and keeps them on the instance;
        ``self.timing`` holds the seconds spent in placement (host),
        rasterisation, solves and copies, and the solver's per-pass kernel
        milliseconds when the environment variable LS_TIMING is set."""
        if np.dtype(self.dtype) != np.dtype(np.float32):
            # the reference loader's ndpointer accepts float32 only (lippmann_schwinger.py:292-304)
            raise TypeError("generate() needs dtype float32, as the reference's solver loader does")
        gs = self.grid_spec
        shape = (gs.nx, gs.ny, gs.nz)
        N = gs.nx * gs.ny * gs.nz
        per_rhs = (9 * (depth + 1) + 24) * N * 4
        k = max(1, min(self.n_samples, int(budget_bytes // (6 * per_rhs)), max_batch or self.n_samples))
        dev = torch.device(self.device)
        self.timing = {"placement_s": 0.0, "rasterise_s": 0.0, "solve_s": 0.0, "copy_s": 0.0, "wait_s": 0.0,
                       "batches": 0, "per_call": k, "solver_ms": {}, "solver_launches": {},
                       "base_reevaluations": 0}
        self._lame_parameters = np.empty((self.n_samples, 2) + shape, dtype=self.dtype)
        self._sigma_bar = np.empty((self.n_samples, 6, 6), dtype=self.dtype)
        self.iterations = np.zeros((self.n_samples, 6), dtype=np.int32)
        fields = torch.empty((k, 2) + shape, dtype=torch.float32, device=dev)
        stage = torch.empty((k, 2) + shape, dtype=torch.float32, pin_memory=True)
        stage_np = stage.numpy()
        solver = LSSolver(gs, nrhs=6 * k, nmat=k, depth=depth)
        eps = torch.empty((6 * k, 6) + shape, dtype=torch.float32, device=dev)
        starts = list(range(0, self.n_samples, k))

        # Host work beside the GPU: while batch b solves (the solver call
        # releases the GIL), one worker thread copies batch b's fields from the
        # pinned staging buffer into the output array and then places the
        # fibres of batch b + 1 (in sample order, so the random streams are
        # those of a sequential run).
        def host_work(copy, n_next):
            t0 = time.perf_counter()
            if copy is not None:
                done, c0, cnt = copy
                done.synchronize()
                self._lame_parameters[c0:c0 + cnt] = stage_np[:cnt]
            t1 = time.perf_counter()
            placed = self.place_samples(n_next, threads=placement_threads)
            return placed, t1 - t0, time.perf_counter() - t1

        pool = ThreadPoolExecutor(max_workers=1)
        try:
            job = pool.submit(host_work, None, min(k, self.n_samples))
            for b, c0 in enumerate(starts):
                cnt = min(k, self.n_samples - c0)
                t0 = time.perf_counter()
                placed, t_copy, t_place = job.result()
                self.timing["wait_s"] += time.perf_counter() - t0
                self.timing["copy_s"] += t_copy
                self.timing["placement_s"] += t_place
                t0 = time.perf_counter()
                for q, (_, pos, rad, ori) in enumerate(placed):
                    self.material_properties_device(pos, rad, ori, out=fields[q])
                    if self.verbose:
                        print(f"sample {c0 + q + 1}/{self.n_samples}: {sum(len(r) for r in rad)} fibres", flush=True)
                stage[:cnt].copy_(fields[:cnt], non_blocking=True)
                done = torch.cuda.Event()
                done.record()
                self.timing["rasterise_s"] += time.perf_counter() - t0
                mu, lam = fields[:cnt, 0], fields[:cnt, 1]
                if swap_lame:
                    lam, mu = mu, lam
                eb = np.tile(np.eye(6, dtype=np.float32), (cnt, 1))
                if cnt < k:
                    # unused slots repeat the last sample
                    pad = k - cnt
                    lam = torch.cat([lam, lam[-1:].expand((pad,) + shape)])
                    mu = torch.cat([mu, mu[-1:].expand((pad,) + shape)])
                    eb = np.tile(np.eye(6, dtype=np.float32), (k, 1))
                mats = np.repeat(np.arange(k, dtype=np.int32), 6)
                lam, mu = lam.contiguous(), mu.contiguous()
                n_next = min(k, self.n_samples - starts[b + 1]) if b + 1 < len(starts) else 0
                job = pool.submit(host_work, (done, c0, cnt), n_next)
                t0 = time.perf_counter()
                res = solver.solve(lam, mu, eb, rhs_mat=mats, epsilon=eps, rtol=rtol, atol=atol, maxiter=maxiter,
                                   check=True)
                self.timing["solve_s"] += time.perf_counter() - t0
                self.timing["batches"] += 1
                ms, cnt_k, nreb = solver.last_timing()
                for kk in ms:
                    self.timing["solver_ms"][kk] = self.timing["solver_ms"].get(kk, 0.0) + ms[kk]
                    self.timing["solver_launches"][kk] = self.timing["solver_launches"].get(kk, 0) + cnt_k[kk]
                self.timing["base_reevaluations"] += nreb
                self._sigma_bar[c0:c0 + cnt] = res["sigma_mean"][:6 * cnt].reshape(cnt, 6, 6)
                self.iterations[c0:c0 + cnt] = res["iters"][:6 * cnt].reshape(cnt, 6)
            t0 = time.perf_counter()
            _, t_copy, _ = job.result()
            self.timing["wait_s"] += time.perf_counter() - t0
            self.timing["copy_s"] += t_copy
        finally:
            pool.shutdown(wait=True)
            solver.close()
        return self._lame_parameters, self._sigma_bar
