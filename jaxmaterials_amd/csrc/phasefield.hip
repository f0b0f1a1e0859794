// Pointwise pieces of the phase-field fracture model on the GPU:
// compute_sigma_damaged (src/jaxmaterials/solver/lippmann_schwinger.py:36-74)
// and compute_strain_energy (:77-100) with the history update of the
// staggered scheme (H = max(H, psi+), tests/phasefield.py:221-223), FP64
// arithmetic on FP32 strain fields (the split is in damage.hpp).
#include <hip/hip_runtime.h>

#include <cstdio>

#include "../../include/lippmann_schwinger.h"
#include "damage.hpp"

namespace {

constexpr int kT = 256;

__global__ void __launch_bounds__(kT) k_damaged_stress(long long N, const float* __restrict__ lam,
                                                       const float* __restrict__ mu, const float* __restrict__ eps,
                                                       const float* __restrict__ dmg, double k,
                                                       float* __restrict__ sigma) {
  const long long st = (long long)gridDim.x * blockDim.x;
  for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < N; v += st) {
    double e[6], sp[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) e[c] = eps[c * N + v];
    const double l = lam[v], m = mu[v];
    const double tr = e[0] + e[1] + e[2];
    lsk::sigma_tension(e, l, m, sp);
    const double gm1 = dmg ? lsk::degradation(dmg[v], k) - 1.0 : 0.0;
#pragma unroll
    for (int c = 0; c < 6; ++c) sigma[c * N + v] = (float)(2.0 * m * e[c] + (c < 3 ? l * tr : 0.0) + gm1 * sp[c]);
  }
}

__global__ void __launch_bounds__(kT) k_strain_energy(long long N, const float* __restrict__ lam,
                                                      const float* __restrict__ mu, const float* __restrict__ eps,
                                                      double* __restrict__ psi, double* __restrict__ hist) {
  const long long st = (long long)gridDim.x * blockDim.x;
  for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < N; v += st) {
    double e[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) e[c] = eps[c * N + v];
    const double p = lsk::strain_energy_tension(e, lam[v], mu[v]);
    if (psi) psi[v] = p;
    if (hist) hist[v] = fmax(hist[v], p);
  }
}

unsigned grid_for(long long N) {
  const long long b = (N + kT - 1) / kT;
  return (unsigned)(b < 4096 ? b : 4096);
}

}  // namespace

extern "C" {

int ls_damaged_stress(const int* voxels, const float* d_lambda, const float* d_mu, const float* d_eps,
                      const float* d_damage, float k, float* d_sigma, void* stream) {
  if (!voxels || !d_lambda || !d_mu || !d_eps || !d_sigma || voxels[0] < 1 || voxels[1] < 1 || voxels[2] < 1) {
    fprintf(stderr, "liblippmannschwinger: ls_damaged_stress: invalid arguments\n");
    return -1;
  }
  const long long N = (long long)voxels[0] * voxels[1] * voxels[2];
  k_damaged_stress<<<grid_for(N), kT, 0, (hipStream_t)stream>>>(N, d_lambda, d_mu, d_eps, d_damage, (double)k,
                                                                 d_sigma);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize((hipStream_t)stream) != hipSuccess) {
    fprintf(stderr, "liblippmannschwinger: ls_damaged_stress failed\n");
    return -1;
  }
  return 0;
}

int ls_strain_energy(const int* voxels, const float* d_lambda, const float* d_mu, const float* d_eps,
                     double* d_psi, double* d_history, void* stream) {
  if (!voxels || !d_lambda || !d_mu || !d_eps || (!d_psi && !d_history) || voxels[0] < 1 || voxels[1] < 1 ||
      voxels[2] < 1) {
    fprintf(stderr, "liblippmannschwinger: ls_strain_energy: invalid arguments\n");
    return -1;
  }
  const long long N = (long long)voxels[0] * voxels[1] * voxels[2];
  k_strain_energy<<<grid_for(N), kT, 0, (hipStream_t)stream>>>(N, d_lambda, d_mu, d_eps, d_psi, d_history);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize((hipStream_t)stream) != hipSuccess) {
    fprintf(stderr, "liblippmannschwinger: ls_strain_energy failed\n");
    return -1;
  }
  return 0;
}

}  // extern "C"
