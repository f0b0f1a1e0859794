"""GPU rasteriser of the reference's layered-fibre generator and the batched
dataset generation, against the oracle (bit-exact fields; C_eff of the FP64
oracle solve within 1e-5 relative, iterations within 1)."""

import numpy as np
import pytest
import torch

from jaxmaterials_amd.common import GridSpec
from oracle import fibres_oracle as fo
from oracle import ls_oracle

pytestmark = pytest.mark.gpu

MU3 = (0.1, 0.5, 1.0)  # (void, fibre, matrix) of the reference's dataset script
LAM3 = (1.0, 0.1, 0.5)


def _generator(gs, n_samples=2, seed=5713853, nlayers=3, d_void=0.01, dtype=np.float32):
    from jaxmaterials_amd.data.fibres import FibreRadiusDistribution, LayeredFibresDatasetGenerator

    rd = FibreRadiusDistribution(r_avg=7.5e-3, r_min=5.0e-3, r_max=10.0e-3, sigma=0.5e-3, gaussian=True)
    return LayeredFibresDatasetGenerator(gs, nlayers, d_void, n_samples, rd, volume_fraction=0.3,
                                         mu_fibre=MU3[1], mu_material=MU3[2], mu_void=MU3[0],
                                         lambda_fibre=LAM3[1], lambda_material=LAM3[2], lambda_void=LAM3[0],
                                         dtype=dtype, rng=np.random.default_rng(seed))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("shape,extents", [((40, 40, 16), (0.2, 0.2, 0.2)), ((160, 160, 64), (0.5, 0.5, 0.2)),
                                           ((33, 30, 14), (0.2, 0.18, 0.2))])
def test_rasteriser_bit_identical_to_oracle(gpu, dtype, shape, extents):
    gs = GridSpec(*shape, *extents)
    gen = _generator(gs, dtype=dtype)
    _, pos, rad, ori = gen.generate_fibre_positions()
    print(f"rasterise {shape} {np.dtype(dtype).name}: {sum(len(r) for r in rad)} fibres", flush=True)
    got = gen.material_properties(pos, rad, ori)
    want = fo.material_properties(gs, gen.d_void, pos, rad, ori, MU3, LAM3, dtype=dtype)
    assert got.dtype == want.dtype and got.shape == want.shape
    assert np.array_equal(got, want)
    assert len(np.unique(got[0])) >= 3


def test_rasteriser_hand_checked_and_edge_cases(gpu):
    from jaxmaterials_amd.data.fibres import rasterise_layered_fibres

    gs = GridSpec(8, 8, 8, 8.0, 8.0, 8.0)
    # no fibres: matrix with void slabs only
    f = rasterise_layered_fibres(gs, 1.0, [], [], [], MU3, LAM3).cpu().numpy()
    assert np.all(f[0][:, :, 1:7] == np.float32(1.0)) and np.all(f[0][:, :, [0, 7]] == np.float32(0.1))
    # one fibre along y through voxel centres (x, z) = (3.5, 3.5), radius 1.01:
    # five voxels per y line; a coincident copy adds its term again
    for copies in (1, 2):
        pos = [np.array([[3.5, 3.5]] * copies)]
        rad = [np.array([1.01] * copies)]
        f = rasterise_layered_fibres(gs, 1.0, pos, rad, [0], MU3, LAM3, dtype=np.float64).cpu().numpy()
        want = fo.material_properties(gs, 1.0, pos, rad, [0], MU3, LAM3, dtype=np.float64)
        assert np.array_equal(f, want)
        assert int((f[0] != f[0, 0, 0, 3]).sum()) == 5 * 8 + 2 * 64
    # orientation 1 tests y
    f = rasterise_layered_fibres(gs, 0.0, [np.array([[2.5, 5.5]])], [np.array([0.6])], [1], MU3, LAM3)
    f = f.cpu().numpy()
    assert np.all(f[0][:, 2, 5] == np.float32(np.float32(1.0) + np.float64(0.5 - 1.0)))
    assert int((f[0] != np.float32(1.0)).sum()) == 8
    with pytest.raises(TypeError):
        rasterise_layered_fibres(gs, 1.0, [], [], [], MU3, LAM3, dtype=np.int32)


def test_generate_matches_oracle_solve(gpu):
    gs = GridSpec(32, 32, 16, 0.2, 0.2, 0.2)
    gen = _generator(gs, n_samples=3, seed=11)
    ref = _generator(gs, n_samples=3, seed=11)
    lame, sbar = gen.generate(max_batch=2)  # two solver calls: a full batch and a padded one
    assert lame.shape == (3, 2, 32, 32, 16) and sbar.shape == (3, 6, 6)
    for s in range(3):
        _, pos, rad, ori = ref.generate_fibre_positions()
        fields = fo.material_properties(gs, ref.d_void, pos, rad, ori, MU3, LAM3, dtype=np.float32)
        assert np.array_equal(lame[s], fields)
        # the reference's loader hands mu to the solver's lambda slot and lambda to its mu slot
        C, its = ls_oracle.effective_tensor(fields[0].astype(np.float64), fields[1].astype(np.float64), gs,
                                            rtol=1e-4, atol=1e-4, maxiter=100)
        # rows whose load case took the same number of updates hold the same
        # iterate up to float32 rounding; one update more or less moves a row
        # by about the stopping residual (1e-4)
        same = np.array([int(a) == int(b) for a, b in zip(gen.iterations[s], its)])
        err = np.abs(sbar[s] - C).max(axis=1) / np.abs(C).max()
        print(f"sample {s}: iterations gpu {gen.iterations[s].tolist()} oracle {its}, "
              f"C_eff row rel err {np.array2string(err, precision=2)}", flush=True)
        assert np.all(err[same] < 1e-5) and np.all(err < 5e-4)
        assert all(abs(int(a) - int(b)) <= 1 for a, b in zip(gen.iterations[s], its))
        assert same.sum() >= 3
