// y and x passes of the iteration with register-resident two-stage FFTs.
//
// For an axis of length N = A * B (A, B compile-time), a line is transformed
// as one Cooley-Tukey split: stage 1 does, per (t < B), the A-point DFT over
// j of the elements t + B j straight from HBM into registers and multiplies
// by W_N^(t k1); the values go through LDS once; stage 2 does, per k1, the
// B-point DFT over t, giving X[k1 + A k2].  The inverse runs the same two
// stages backwards.  Against the generic LDS Stockham engine (fft_lds.hpp:
// one LDS round trip per radix stage plus the load and store staging, and
// table twiddles inside every butterfly) a transform costs one LDS round trip
// and its inner twiddles are constants.
//
// LDS layout of a block's lines: element pos of line l at
// l * Np + pos + pos / B with Np = N + N / B + 1 (odd).  Stage 1 and the
// stores walk consecutive lines across lanes (2 Np is coprime to the 64 banks'
// pairing), stage 2 of the x pass walks consecutive k1 (stride B + 1, odd).
//
// The x pass keeps the six Voigt components of a mode in one thread for stage
// 2, so the divergence, residual and Gamma0 act on registers between the
// forward and inverse stage 2 and the block needs two barriers in all.
#pragma once
#include "fft_reg.hpp"
#include "ls_kernels.hpp"

namespace lsk {

template <int N, int B>
struct RegLayout {
  static constexpr int Np = N + N / B + 1;
  static __device__ __forceinline__ int at(int l, int pos) { return l * Np + pos + pos / B; }
};

// ---------------------------------------------------------------------------
// y pass: FFT along y (ny = A B) of T consecutive (kz, component) columns at
// fixed x of one load case (same tiling as k_ypass).
template <int A, int B, int T, int NT, bool INV>
__global__ void __launch_bounds__(NT) k_ypass_r(GridDev g, const RhsState* __restrict__ st,
                                                float2* __restrict__ spec, int ntiles) {
  constexpr int N = A * B;
  using L = RegLayout<N, B>;
  extern __shared__ float2 lds[];
  const int lin = xcd_remap(blockIdx.x, gridDim.x);
  const int per_rhs = g.nx * ntiles;
  const int r = lin / per_rhs;
  const int tile = lin - r * per_rhs;
  if (!st[r].active) return;
  const int Q = 6 * g.P;  // columns per y row
  const int x = tile / ntiles;
  const int q0 = (tile - x * ntiles) * T;
  const int nq = min(T, Q - q0);
  float2* S = spec + (long long)r * g.M * 6 + (long long)x * N * Q + q0;
  const float2* __restrict__ tw = g.py.tw;

  for (int it = threadIdx.x; it < B * T; it += NT) {
    const int t = it / T, l = it - t * T;
    float2 v[A];
#pragma unroll
    for (int j = 0; j < A; ++j) v[j] = (l < nq) ? S[(long long)(t + B * j) * Q + l] : make_float2(0.f, 0.f);
    dft_reg<A, INV>(v);
#pragma unroll
    for (int k1 = 1; k1 < A; ++k1) v[k1] = INV ? cmulc(v[k1], tw[t * k1]) : cmul(v[k1], tw[t * k1]);
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) lds[L::at(l, t + B * k1)] = v[k1];
  }
  __syncthreads();
  for (int it = threadIdx.x; it < A * T; it += NT) {
    const int k1 = it / T, l = it - k1 * T;
    float2 u[B];
#pragma unroll
    for (int t = 0; t < B; ++t) u[t] = lds[L::at(l, t + B * k1)];
    dft_reg<B, INV>(u);
    if (l < nq) {
#pragma unroll
      for (int k2 = 0; k2 < B; ++k2) S[(long long)(k1 + A * k2) * Q + l] = u[k2];
    }
  }
}

// ---------------------------------------------------------------------------
// x pass (nx = A B): forward FFT along x -> d = d_base + sigma_hat . xi~ ->
// residual partial -> Gamma0(d) -> inverse FFT along x, for TM consecutive
// (ky, kz) modes of one load case (all six components).  Semantics as
// k_xpass (lippmann_schwinger.py:118-143, fourier.py:115-186,
// fourier.cu:140-327).
template <int A, int B, int TM, int NT>
__global__ void __launch_bounds__(NT) k_xpass_r(GridDev g, const RhsDesc* __restrict__ rhs,
                                                const RhsState* __restrict__ st, float2* __restrict__ spec,
                                                const float2* __restrict__ dbase, double* __restrict__ partials,
                                                float* __restrict__ dc, int ntiles, long long db_per_rhs,
                                                int do_gamma) {
  constexpr int N = A * B;
  constexpr int LM = 6 * TM;
  using L = RegLayout<N, B>;
  extern __shared__ float2 lds[];
  __shared__ double red[NT / 64];
  const int lin = xcd_remap(blockIdx.x, gridDim.x);
  const int r = lin / ntiles;
  const int tile = lin - r * ntiles;
  if (!st[r].active) return;
  const int P = g.P;
  const int nm = g.ny * P;  // modes per x plane
  const int m0 = tile * TM;
  const int nq = min(LM, 6 * (nm - m0));
  const long long colmax = 6LL * nm;
  float2* S = spec + (long long)r * g.M * 6 + 6LL * m0;
  // d_base tile of this block: [x][TM][3] (DbLayout with Tm = TM)
  const float2* DB = dbase ? dbase + (long long)r * db_per_rhs + (long long)tile * N * TM * 3 : nullptr;
  const float2* __restrict__ tw = g.px.tw;
  const RhsDesc d = rhs[r];

  // stage 1 (forward): item (t, l)
  for (int it = threadIdx.x; it < B * LM; it += NT) {
    const int t = it / LM, l = it - t * LM;
    float2 v[A];
#pragma unroll
    for (int j = 0; j < A; ++j) v[j] = (l < nq) ? S[(t + B * j) * colmax + l] : make_float2(0.f, 0.f);
    dft_reg<A, false>(v);
#pragma unroll
    for (int k1 = 1; k1 < A; ++k1) v[k1] = cmul(v[k1], tw[t * k1]);
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) lds[L::at(l, t + B * k1)] = v[k1];
  }
  __syncthreads();

  // stage 2: item (k1, m): forward B-point DFTs of the six components, the
  // mode-local work on the B wavenumbers kx = k1 + A k2, inverse B-point DFTs
  double acc = 0.0;
  for (int it = threadIdx.x; it < A * TM; it += NT) {
    const int m = it / A, k1 = it - m * A;
    const int gm = m0 + m;
    if (gm >= nm) continue;
    const int ky = fdiv(gm, g.f_P), kz = gm - ky * P;
    float2 u[6][B];
#pragma unroll
    for (int c = 0; c < 6; ++c) {
#pragma unroll
      for (int t = 0; t < B; ++t) u[c][t] = lds[L::at(6 * m + c, t + B * k1)];
      dft_reg<B, false>(u[c]);
    }
    const float syk = g.sy[ky], cyk = g.cy[ky];
    const float szk = g.sz[kz], czk = g.cz[kz];
    const double wk = (kz == 0 || 2 * kz == g.nz) ? 1.0 : 2.0;
    double e2 = 0.0;
#pragma unroll
    for (int k2 = 0; k2 < B; ++k2) {
      const int kx = k1 + A * k2;
      if (gm == 0 && kx == 0) {
#pragma unroll
        for (int c = 0; c < 6; ++c) dc[r * 6 + c] = u[c][0].x;
      }
      const float sxk = g.sx[kx], cxk = g.cx[kx];
      const float k0 = g.gx * sxk * cyk * czk;
      const float kk1 = g.gy * cxk * syk * czk;
      const float kk2 = g.gz * cxk * cyk * szk;
      float2 d0, d1, d2;
      d0.x = k0 * u[0][k2].x + kk1 * u[3][k2].x + kk2 * u[4][k2].x;
      d0.y = k0 * u[0][k2].y + kk1 * u[3][k2].y + kk2 * u[4][k2].y;
      d1.x = k0 * u[3][k2].x + kk1 * u[1][k2].x + kk2 * u[5][k2].x;
      d1.y = k0 * u[3][k2].y + kk1 * u[1][k2].y + kk2 * u[5][k2].y;
      d2.x = k0 * u[4][k2].x + kk1 * u[5][k2].x + kk2 * u[2][k2].x;
      d2.y = k0 * u[4][k2].y + kk1 * u[5][k2].y + kk2 * u[2][k2].y;
      if (DB) {
        const float2* b = DB + (kx * TM + m) * 3;
        const float2 b0 = b[0], b1 = b[1], b2 = b[2];
        d0.x += b0.x;
        d0.y += b0.y;
        d1.x += b1.x;
        d1.y += b1.y;
        d2.x += b2.x;
        d2.y += b2.y;
      }
      e2 += (double)d0.x * d0.x + (double)d0.y * d0.y + (double)d1.x * d1.x + (double)d1.y * d1.y +
            (double)d2.x * d2.x + (double)d2.y * d2.y;
      if (do_gamma) {
        const float q2 = k0 * k0 + kk1 * kk1 + kk2 * kk2;
        if (q2 > 0.f) {
          const float inv = 1.0f / sqrtf(q2);
          const float n0 = k0 * inv, n1 = kk1 * inv, n2 = kk2 * inv;
          const float2 v0 = cscale(d0, inv), v1 = cscale(d1, inv), v2 = cscale(d2, inv);
          const float2 rho = make_float2(n0 * v0.x + n1 * v1.x + n2 * v2.x, n0 * v0.y + n1 * v1.y + n2 * v2.y);
          const float cA = d.cA, cB = d.cB;
          const float h = 0.5f * cA;
          u[0][k2] = make_float2(cA * n0 * v0.x + cB * rho.x * n0 * n0, cA * n0 * v0.y + cB * rho.y * n0 * n0);
          u[1][k2] = make_float2(cA * n1 * v1.x + cB * rho.x * n1 * n1, cA * n1 * v1.y + cB * rho.y * n1 * n1);
          u[2][k2] = make_float2(cA * n2 * v2.x + cB * rho.x * n2 * n2, cA * n2 * v2.y + cB * rho.y * n2 * n2);
          u[3][k2] = make_float2(h * (n0 * v1.x + n1 * v0.x) + cB * rho.x * n0 * n1,
                                 h * (n0 * v1.y + n1 * v0.y) + cB * rho.y * n0 * n1);
          u[4][k2] = make_float2(h * (n0 * v2.x + n2 * v0.x) + cB * rho.x * n0 * n2,
                                 h * (n0 * v2.y + n2 * v0.y) + cB * rho.y * n0 * n2);
          u[5][k2] = make_float2(h * (n1 * v2.x + n2 * v1.x) + cB * rho.x * n1 * n2,
                                 h * (n1 * v2.y + n2 * v1.y) + cB * rho.y * n1 * n2);
        } else {
#pragma unroll
          for (int c = 0; c < 6; ++c) u[c][k2] = make_float2(0.f, 0.f);
        }
      }
    }
    acc += wk * e2;
    if (do_gamma) {
      float2 twk[B];
#pragma unroll
      for (int t = 1; t < B; ++t) twk[t] = tw[t * k1];
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        dft_reg<B, true>(u[c]);
#pragma unroll
        for (int t = 1; t < B; ++t) u[c][t] = cmulc(u[c][t], twk[t]);
#pragma unroll
        for (int t = 0; t < B; ++t) lds[L::at(6 * m + c, t + B * k1)] = u[c][t];
      }
    }
  }
  const double bs = block_sum<NT>(acc, red);
  if (threadIdx.x == 0) partials[(long long)r * ntiles + tile] = bs;
  if (!do_gamma) return;
  __syncthreads();

  // stage 1 (inverse): item (t, l)
  for (int it = threadIdx.x; it < B * LM; it += NT) {
    const int t = it / LM, l = it - t * LM;
    float2 v[A];
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) v[k1] = lds[L::at(l, t + B * k1)];
    dft_reg<A, true>(v);
    if (l < nq) {
#pragma unroll
      for (int j = 0; j < A; ++j) S[(t + B * j) * colmax + l] = v[j];
    }
  }
}

}  // namespace lsk
