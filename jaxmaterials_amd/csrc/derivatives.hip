// Real-space backward divergence and relative divergence on the GPU.
//
// D^-_d g = (n_d / (4 L_d)) (1 + S_e)(1 + S_f)(1 - S_d) g on the rotated
// staggered grid, S_d the periodic shift g(i) -> g(i - 1) along axis d and
// e, f the other two axes (src/jaxmaterials/solver/derivatives.py:8-63,
// cuda/src/derivatives.cu:9-124): the difference along d of the 2 x 2
// average over the other two axes, taken between voxel i and i - 1.  The
// divergence of a Voigt stress field is row-wise (derivatives.py:66-96), and
// the relative divergence ||D^- . sigma|| / (sqrt(N) ||<sigma>||_V) is the
// reference's real-space stopping residual (lippmann_schwinger.py:103-115),
// equal to the Fourier residual by Parseval.  This is an independent check
// of the iteration's residual; the iteration itself never calls it.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>

#include "../../include/lippmann_schwinger.h"

namespace {

constexpr int kThreads = 256;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Per voxel: the three divergence components from the 2 x 2 x 2 cell of
// voxels (i - a, j - b, k - c), a, b, c in {0, 1}; FP64 block partials of
// sum |div|^2 and of the six stress components.  part[block][7].
__global__ void __launch_bounds__(kThreads) k_divergence(int nx, int ny, int nz, float sx, float sy, float sz,
                                                         const float* __restrict__ sigma, float* __restrict__ div,
                                                         double* __restrict__ part) {
  __shared__ double red[7][kThreads / 64];
  const long long N = (long long)nx * ny * nz;
  const long long v = (long long)blockIdx.x * kThreads + threadIdx.x;
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  if (v < N) {
    const int k = (int)(v % nz);
    const long long q = v / nz;
    const int j = (int)(q % ny);
    const int i = (int)(q / ny);
    const int im = i == 0 ? nx - 1 : i - 1, jm = j == 0 ? ny - 1 : j - 1, km = k == 0 ? nz - 1 : k - 1;
    const long long ii[2] = {(long long)i * ny * nz, (long long)im * ny * nz};
    const long long jj[2] = {(long long)j * nz, (long long)jm * nz};
    const int kk[2] = {k, km};
    // c[x][y][z] of one component at (i - x, j - y, k - z)
    auto cell = [&](int comp, float c[2][2][2]) {
      const float* s = sigma + (long long)comp * N;
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
          for (int e = 0; e < 2; ++e) c[a][b][e] = s[ii[a] + jj[b] + kk[e]];
    };
    // D^-_x, D^-_y, D^-_z of a cell
    auto dx = [&](const float c[2][2][2]) {
      float t = 0.f;
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int e = 0; e < 2; ++e) t += c[0][b][e] - c[1][b][e];
      return sx * t;
    };
    auto dy = [&](const float c[2][2][2]) {
      float t = 0.f;
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int e = 0; e < 2; ++e) t += c[a][0][e] - c[a][1][e];
      return sy * t;
    };
    auto dz = [&](const float c[2][2][2]) {
      float t = 0.f;
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) t += c[a][b][0] - c[a][b][1];
      return sz * t;
    };
    float cxx[2][2][2], cyy[2][2][2], czz[2][2][2], cxy[2][2][2], cxz[2][2][2], cyz[2][2][2];
    cell(0, cxx);
    cell(1, cyy);
    cell(2, czz);
    cell(3, cxy);
    cell(4, cxz);
    cell(5, cyz);
    const float d0 = dx(cxx) + dy(cxy) + dz(cxz);
    const float d1 = dx(cxy) + dy(cyy) + dz(cyz);
    const float d2 = dx(cxz) + dy(cyz) + dz(czz);
    if (div) {
      div[v] = d0;
      div[N + v] = d1;
      div[2 * N + v] = d2;
    }
    acc[0] = (double)d0 * d0 + (double)d1 * d1 + (double)d2 * d2;
    acc[1] = cxx[0][0][0];
    acc[2] = cyy[0][0][0];
    acc[3] = czz[0][0][0];
    acc[4] = cxy[0][0][0];
    acc[5] = cxz[0][0][0];
    acc[6] = cyz[0][0][0];
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < 7; ++c) {
    const double s = wave_sum(acc[c]);
    if (lane == 0) red[c][w] = s;
  }
  __syncthreads();
  if (threadIdx.x < 7) {
    double s = 0.0;
    for (int q = 0; q < kThreads / 64; ++q) s += red[threadIdx.x][q];
    part[(long long)blockIdx.x * 7 + threadIdx.x] = s;
  }
}

// Sum of the block partials in a fixed order: out[7].
__global__ void __launch_bounds__(kThreads) k_sum7(const double* __restrict__ part, long long nblk,
                                                  double* __restrict__ out) {
  __shared__ double red[kThreads / 64];
  for (int c = 0; c < 7; ++c) {
    double a = 0.0;
    for (long long b = threadIdx.x; b < nblk; b += kThreads) a += part[b * 7 + c];
    a = wave_sum(a);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
      double s = 0.0;
      for (int q = 0; q < kThreads / 64; ++q) s += red[q];
      out[c] = s;
    }
    __syncthreads();
  }
}

}  // namespace

extern "C" int ls_relative_divergence(const int* voxels, const float* extents, const float* d_sigma, float* d_div,
                                      double* rel, void* stream) {
  if (!voxels || !extents || !d_sigma || !rel) return -1;
  const int nx = voxels[0], ny = voxels[1], nz = voxels[2];
  if (nx < 1 || ny < 1 || nz < 1 || !(extents[0] > 0.f && extents[1] > 0.f && extents[2] > 0.f)) {
    fprintf(stderr, "liblippmannschwinger: ls_relative_divergence: invalid grid\n");
    return -1;
  }
  const long long N = (long long)nx * ny * nz;
  const long long nblk = (N + kThreads - 1) / kThreads;
  hipStream_t st = (hipStream_t)stream;
  double* part = nullptr;
  double host[7];
  int rc = -1;
  do {
    if (hipMalloc((void**)&part, (size_t)(nblk * 7 + 7) * sizeof(double)) != hipSuccess) break;
    // (n_d / (4 L_d)) as in derivatives.py:8-63 (extents cross the ABI as float)
    const float sx = (float)(0.25 * nx / (double)extents[0]);
    const float sy = (float)(0.25 * ny / (double)extents[1]);
    const float sz = (float)(0.25 * nz / (double)extents[2]);
    k_divergence<<<(unsigned)nblk, kThreads, 0, st>>>(nx, ny, nz, sx, sy, sz, d_sigma, d_div, part);
    k_sum7<<<1, kThreads, 0, st>>>(part, nblk, part + nblk * 7);
    if (hipGetLastError() != hipSuccess) break;
    if (hipMemcpyAsync(host, part + nblk * 7, sizeof(host), hipMemcpyDeviceToHost, st) != hipSuccess) break;
    if (hipStreamSynchronize(st) != hipSuccess) break;
    double den = 0.0;
    for (int c = 0; c < 6; ++c) {
      const double m = host[1 + c] / (double)N;
      den += (c < 3 ? 1.0 : 2.0) * m * m;
    }
    *rel = sqrt(host[0]) / (sqrt((double)N) * sqrt(den));
    rc = 0;
  } while (0);
  if (part) (void)hipFree(part);
  if (rc != 0) fprintf(stderr, "liblippmannschwinger: ls_relative_divergence failed\n");
  return rc;
}
