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
#pragma once
#include <type_traits>

#include "fft_reg.hpp"
#include "ls_kernels.hpp"

namespace lsk {

template <int N, int B>
struct RegLayout {
  static constexpr int Np = N + N / B + 1;
  static __device__ __forceinline__ int at(int l, int pos) { return l * Np + pos + pos / B; }
};

// The same accesses without padding: within each aligned group of 32
// positions of line l, position pos is stored at pos ^ ((l ^ (pos / 32)) & 31)
// restricted so that (pos / 32) only flips the low log2(B) bits.  With
// pos = t + B k1: the 32 lanes of a stage-2 access (k1 varying) hit 32
// different banks, and the lanes of a stage-1/3 access (lines l varying)
// spread over the banks as with the padded layout, in N complex per line
// instead of N + N / B + 1.  Needs N % 32 == 0 and B a power of two <= 32.
template <int N, int B>
struct SwzLayout {
  static_assert(N % 32 == 0 && (B & (B - 1)) == 0 && B <= 32, "swizzled layout needs N % 32 == 0, B = 2^k <= 32");
  static __device__ __forceinline__ int at(int l, int pos) {
    return l * N + (pos ^ ((l ^ ((pos >> 5) & (B - 1))) & 31));
  }
};

template <int N, int B>
constexpr bool swz_ok() {
  return N % 32 == 0 && (B & (B - 1)) == 0 && B <= 32;
}

// ---------------------------------------------------------------------------
// y pass: FFT along y (ny = A B) of T consecutive (kz, component) columns at
// fixed x of one load case (same tiling as k_ypass).
template <int A, int B, int T, int NT, bool INV>
__global__ void __launch_bounds__(NT) k_ypass_r(GridDev g, const RhsState* __restrict__ st,
                                                float2* __restrict__ spec, int ntiles, AAPlanes pl) {
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
      if constexpr (INV) {
        if (pl.im) {
#pragma unroll
          for (int k2 = 0; k2 < B; ++k2) aa_plane_record(pl, g, r, x, k1 + A * k2, q0 + l, u[k2].y);
        }
      }
    }
  }
}

}  // namespace lsk

namespace lsk {

// ---------------------------------------------------------------------------
// x pass transforming three lines each way instead of six.
//
// With xi~ = (alpha s_x(k), beta c_x(k), gamma c_x(k)), s_x = sin(pi k / nx),
// c_x = cos(pi k / nx) (alpha = gx c_y c_z, beta = gy s_y c_z, gamma = gz c_y s_z
// depend on the mode (ky, kz) only), the divergence of a line is
//   d_i(k) = alpha s_x F[sigma_ix] + c_x F[w_i],  w_i = beta sigma_iy + gamma sigma_iz,
// and with s_x = e^{-i pi k/n} (e^{2 pi i k/n} - 1) / 2i,
//          c_x = e^{-i pi k/n} (e^{2 pi i k/n} + 1) / 2:
//   d_i(k) = e^{-i pi k/n} F[h_i](k),
//   h_i(x) = alpha (sigma_ix(x+1) - sigma_ix(x)) / 2i + (w_i(x) + w_i(x+1)) / 2,
// so the forward FFT along x acts on the three lines h_i.  Gamma0 in terms
// of the divergence is r_hat = sym(xi~ (x) u), u = (cA v + cB rho n) / |xi~|
// (v = d / |xi~|, n = xi~ / |xi~|, rho = n . v): with u = e^{-i pi k/n} u',
//   r_ij(x) = (P_ij + P_ji) / 2,  P_xj = alpha (g_j(x) - g_j(x-1)) / 2i,
//   P_yj = beta (g_j(x) + g_j(x-1)) / 2,  P_zj = gamma (g_j(x) + g_j(x-1)) / 2,
// g = IFFT[u'], so the inverse FFT along x also acts on three lines.  The
// residual |d|^2 = |F[h]|^2 is unchanged; d_base (the FP64 divergence of the
// base stress) enters as e^{+i pi k/n} d_base.  Same results as k_xpass up to
// FP32 rounding (lippmann_schwinger.py:118-143, fourier.py:115-186).
//
// LDS: h and then g lines [x][3 TM] in the first 3 TM N slots, the stage-1
// transforms in RegLayout after them; per-mode coefficients in a small table.
// INPL: the transforms overwrite the h / g lines in place (each thread holds
// at most one line segment of stages 1 and 3 in registers across a barrier),
// which halves the LDS of a block.
// REC: 0 plain; 1 Anderson (record y of the residual); 2 conjugate gradients
// (d = alpha e^{-i pi k/n} F[h] + d, written back to the base-divergence
// buffer, and the block partial of the residual's C0 norm, see CgState).
template <int A, int B, int TM, int NT, int REC, int MODE>
__global__ void __launch_bounds__(NT) k_xpass_h(GridDev g, const RhsDesc* __restrict__ rhs,
                                                const RhsState* __restrict__ st, float2* __restrict__ spec,
                                                float2* __restrict__ dbase, double* __restrict__ partials,
                                                float* __restrict__ dc, int ntiles, long long db_per_rhs,
                                                int do_gamma, AAParams aa) {
  constexpr int N = A * B;
  constexpr int LH = 3 * TM;  // h / g lines
  constexpr bool INPL = MODE != 0;
  constexpr bool SPLIT = MODE == 2;
  static_assert(!SPLIT || swz_ok<N, B>(), "the split stage 2 runs on the swizzled layout");
  // in place: the transforms share the h / g lines' N LH complex, unpadded
  using L = std::conditional_t<INPL && swz_ok<N, B>(), SwzLayout<N, B>, RegLayout<N, B>>;
  static_assert(!INPL || B * LH <= NT, "in-place stages 1 and 3 need one item per thread");
  extern __shared__ float2 lds[];
  float2* H = lds;                       // [x][LH]
  float2* W = INPL ? lds : lds + LH * N;  // RegLayout lines 3m + i
  __shared__ double red[NT / 64];
  // per mode: alpha, beta, gamma (xi~ = (alpha s_x, beta c_x, gamma c_x)),
  // c_y c_z, s_y c_z, c_y s_z and the Hermitian weight
  __shared__ float coef[TM][4];
  __shared__ int valid[TM];
  const int lin = xcd_remap(blockIdx.x, gridDim.x);
  const int r = lin / ntiles;
  const int tile = lin - r * ntiles;
  if (!st[r].active) return;
  const int P = g.P;
  const int nm = g.ny * P;
  const int m0 = tile * TM;
  const long long colmax = 6LL * nm;
  float2* S = spec + (long long)r * g.M * 6 + 6LL * m0;
  float2* DB = dbase ? dbase + (long long)r * db_per_rhs + (long long)tile * N * TM * 3 : nullptr;
  constexpr bool CG = REC == 2;
  // the x twiddles e^{-2 pi i j / N} in LDS (indices t k1 < N of stages 1 and 2)
  __shared__ float2 tw[N];
  for (int i = threadIdx.x; i < N; i += NT) tw[i] = g.px.tw[i];
  const RhsDesc d = rhs[r];
  if (threadIdx.x < TM) {
    const int gm = m0 + threadIdx.x;
    const int ok = gm < nm;
    const int ky = ok ? fdiv(gm, g.f_P) : 0, kz = ok ? gm - ky * P : 0;
    const float cyk = g.cy[ky], syk = g.sy[ky], czk = g.cz[kz], szk = g.sz[kz];
    coef[threadIdx.x][0] = g.gx * cyk * czk;
    coef[threadIdx.x][1] = g.gy * syk * czk;
    coef[threadIdx.x][2] = g.gz * cyk * szk;
    coef[threadIdx.x][3] = (kz == 0 || 2 * kz == g.nz) ? 1.f : 2.f;
    valid[threadIdx.x] = ok;
  }
  __syncthreads();

  // stage 0: h_i(x) from the stress at x and x + 1.
  // h_i = alpha (b_ix - a_ix) / 2i + (w_i(a) + w_i(b)) / 2, w_i = beta s_iy + gamma s_iz,
  // a / b the stress at x / x + 1 (components 0 xx, 1 yy, 2 zz, 3 xy, 4 xz, 5 yz;
  // i = x: (xx, xy, xz); i = y: (xy, yy, yz); i = z: (xz, yz, zz))
  auto h_of = [&](int m, const float4& a01, const float4& a23, const float4& a45, const float4& b01,
                  const float4& b23, const float4& b45, int x) {
    const float ha = 0.5f * coef[m][0], be = coef[m][1], ga = coef[m][2];
    const float wx0r = be * (a23.z + b23.z) + ga * (a45.x + b45.x), wx0i = be * (a23.w + b23.w) + ga * (a45.y + b45.y);
    const float wx1r = be * (a01.z + b01.z) + ga * (a45.z + b45.z), wx1i = be * (a01.w + b01.w) + ga * (a45.w + b45.w);
    const float wx2r = be * (a45.z + b45.z) + ga * (a23.x + b23.x), wx2i = be * (a45.w + b45.w) + ga * (a23.y + b23.y);
    H[x * LH + 3 * m] = make_float2(ha * (b01.y - a01.y) + 0.5f * wx0r, -ha * (b01.x - a01.x) + 0.5f * wx0i);
    H[x * LH + 3 * m + 1] = make_float2(ha * (b23.w - a23.w) + 0.5f * wx1r, -ha * (b23.z - a23.z) + 0.5f * wx1i);
    H[x * LH + 3 * m + 2] = make_float2(ha * (b45.y - a45.y) + 0.5f * wx2r, -ha * (b45.x - a45.x) + 0.5f * wx2i);
  };
  // one item per (x, mode); the row x + 1 is read again by the neighbouring
  // item and comes from the cache
  for (int it = threadIdx.x; it < N * TM; it += NT) {
    const int x = it / TM, m = it - x * TM;
    if (valid[m]) {
      const int x1 = (x + 1 == N) ? 0 : x + 1;
      const float2* p0 = S + x * colmax + 6 * m;
      const float2* p1 = S + x1 * colmax + 6 * m;
      h_of(m, ld4(p0), ld4(p0 + 2), ld4(p0 + 4), ld4(p1), ld4(p1 + 2), ld4(p1 + 4), x);
    } else {
      H[x * LH + 3 * m] = H[x * LH + 3 * m + 1] = H[x * LH + 3 * m + 2] = make_float2(0.f, 0.f);
    }
  }
  if (tile == 0 && threadIdx.x < 64) {
    // mean stress of the line kx = 0 of mode 0 (denominator of the residual):
    // one wave sums the six components over x
    float acc6[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int x = threadIdx.x; x < N; x += 64) {
      const float2* p0 = S + x * colmax;
#pragma unroll
      for (int c = 0; c < 6; ++c) acc6[c] += p0[c].x;
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      double v = wave_sum((double)acc6[c]);
      if (threadIdx.x == 0) dc[r * 6 + c] = (float)v;
    }
  }
  __syncthreads();

  // stage 1 (forward): item (t, line), A-point DFTs over x = t + B j
  if constexpr (INPL) {
    const int it = threadIdx.x;
    const int t = it / LH, l = it - t * LH;
    float2 v[A];
    if (it < B * LH) {
#pragma unroll
      for (int j = 0; j < A; ++j) v[j] = H[(t + B * j) * LH + l];
      dft_reg<A, false>(v);
#pragma unroll
      for (int k1 = 1; k1 < A; ++k1) v[k1] = cmul(v[k1], tw[t * k1]);
    }
    __syncthreads();
    if (it < B * LH) {
#pragma unroll
      for (int k1 = 0; k1 < A; ++k1) W[L::at(l, t + B * k1)] = v[k1];
    }
  } else {
    for (int it = threadIdx.x; it < B * LH; it += NT) {
      const int t = it / LH, l = it - t * LH;
      float2 v[A];
#pragma unroll
      for (int j = 0; j < A; ++j) v[j] = H[(t + B * j) * LH + l];
      dft_reg<A, false>(v);
#pragma unroll
      for (int k1 = 1; k1 < A; ++k1) v[k1] = cmul(v[k1], tw[t * k1]);
#pragma unroll
      for (int k1 = 0; k1 < A; ++k1) W[L::at(l, t + B * k1)] = v[k1];
    }
  }
  __syncthreads();

  // stage 2: item (k1, m): B-point DFTs of h, residual and Gamma0 per kx,
  // inverse B-point DFTs of u' (with Anderson, y of u' is recorded for the Gram row)
  double acc = 0.0;
  const bool rec = REC == 1 && do_gamma;
  [[maybe_unused]] double acc2 = 0.0;
  [[maybe_unused]] const float cga = CG ? (float)aa.cg[r].alpha : 1.f;
  [[maybe_unused]] const float ccg = CG ? (d.lam0 + d.mu0) / (d.lam0 + 2.f * d.mu0) : 0.f;
  float2* Un = rec ? aa.U + ((long long)r * aa.m1 + aa.head[r]) * aa.u_per_rhs + (long long)tile * N * TM * 3
                   : nullptr;
  if constexpr (!SPLIT) {
    for (int it = threadIdx.x; it < A * TM; it += NT) {
      const int m = it / A, k1 = it - m * A;
      if (!valid[m]) continue;
      float2 u[3][B];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
#pragma unroll
        for (int t = 0; t < B; ++t) u[i][t] = W[L::at(3 * m + i, t + B * k1)];
        dft_reg<B, false>(u[i]);
      }
      const float al = coef[m][0], be = coef[m][1], ga = coef[m][2];
      double e2 = 0.0;
#pragma unroll
      for (int k2 = 0; k2 < B; ++k2) {
        const int kx = k1 + A * k2;
        float2 d0 = u[0][k2], d1 = u[1][k2], d2 = u[2][k2];
        // tile entry (kx, m, i) at ((k2 3 + i) TM + m) A + k1 (DbLayout, three-line order)
        const int eo = (k2 * 3 * TM + m) * A + k1;
        if constexpr (CG) {
          d0 = cscale(d0, cga);
          d1 = cscale(d1, cga);
          d2 = cscale(d2, cga);
        }
        if (DB) {
          const float2 ph = g.phx[kx];
          d0 = cadd(d0, cmul(ph, DB[eo]));
          d1 = cadd(d1, cmul(ph, DB[eo + TM * A]));
          d2 = cadd(d2, cmul(ph, DB[eo + 2 * TM * A]));
          if constexpr (CG) {
            DB[eo] = cmulc(d0, ph);
            DB[eo + TM * A] = cmulc(d1, ph);
            DB[eo + 2 * TM * A] = cmulc(d2, ph);
          }
        }
        const double e2m = (double)d0.x * d0.x + (double)d0.y * d0.y + (double)d1.x * d1.x + (double)d1.y * d1.y +
                           (double)d2.x * d2.x + (double)d2.y * d2.y;
        e2 += e2m;
        const float sxk = g.sx[kx], cxk = g.cx[kx];
        const float k0 = al * sxk;
        const float kk1 = be * cxk;
        const float kk2 = ga * cxk;
        const float q2 = k0 * k0 + kk1 * kk1 + kk2 * kk2;
        if (q2 > 0.f) {
          const float inv = 1.0f / sqrtf(q2);
          const float n0 = k0 * inv, n1 = kk1 * inv, n2 = kk2 * inv;
          const float rhox = (n0 * d0.x + n1 * d1.x + n2 * d2.x) * inv, rhoy = (n0 * d0.y + n1 * d1.y + n2 * d2.y) * inv;
          if constexpr (CG)
            acc2 += (double)coef[m][3] *
                    (e2m / (double)q2 - (double)ccg * ((double)rhox * rhox + (double)rhoy * rhoy));
          const float cA = d.cA * inv * inv, cB = d.cB * inv;
          u[0][k2] = make_float2(cA * d0.x + cB * rhox * n0, cA * d0.y + cB * rhoy * n0);
          u[1][k2] = make_float2(cA * d1.x + cB * rhox * n1, cA * d1.y + cB * rhoy * n1);
          u[2][k2] = make_float2(cA * d2.x + cB * rhox * n2, cA * d2.y + cB * rhoy * n2);
          if (rec)
            aa_y(coef[m][3], q2, inv, n0, n1, n2, u[0][k2], u[1][k2], u[2][k2], Un[eo], Un[eo + TM * A],
                 Un[eo + 2 * TM * A]);
        } else {
          u[0][k2] = u[1][k2] = u[2][k2] = make_float2(0.f, 0.f);
          if (rec) Un[eo] = Un[eo + TM * A] = Un[eo + 2 * TM * A] = make_float2(0.f, 0.f);
        }
      }
      acc += (double)coef[m][3] * e2;
      if (do_gamma) {
        float2 twk[B];
#pragma unroll
        for (int t = 1; t < B; ++t) twk[t] = tw[t * k1];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          dft_reg<B, true>(u[i]);
#pragma unroll
          for (int t = 1; t < B; ++t) u[i][t] = cmulc(u[i][t], twk[t]);
#pragma unroll
          for (int t = 0; t < B; ++t) W[L::at(3 * m + i, t + B * k1)] = u[i][t];
        }
      }
    }
  } else {
    // the same work in three steps with one transform line (or one mode
    // point) per item, so that few registers are live: (a) forward B-point
    // DFTs in place, the output k2 of item (k1, line) at position k2 + B k1;
    // (b) residual, d_base, Gamma0 and the Anderson record per (kx, mode);
    // (c) inverse B-point DFTs and the conjugate twiddles, in place.
    for (int it = threadIdx.x; it < A * LH; it += NT) {
      const int line = it / A, k1 = it - line * A;
      if (!valid[line / 3]) continue;
      float2 u[B];
#pragma unroll
      for (int t = 0; t < B; ++t) u[t] = W[L::at(line, t + B * k1)];
      dft_reg<B, false>(u);
#pragma unroll
      for (int k2 = 0; k2 < B; ++k2) W[L::at(line, k2 + B * k1)] = u[k2];
    }
    __syncthreads();
    for (int it = threadIdx.x; it < N * TM; it += NT) {
      const int m = it / N, kx = it - m * N;
      if (!valid[m]) continue;
      const int k2 = kx / A, k1 = kx - k2 * A;
      const int p0 = L::at(3 * m, k2 + B * k1), p1 = L::at(3 * m + 1, k2 + B * k1),
                p2 = L::at(3 * m + 2, k2 + B * k1);
      float2 d0 = W[p0], d1 = W[p1], d2 = W[p2];
      const int eo = (k2 * 3 * TM + m) * A + k1;
      if constexpr (CG) {
        d0 = cscale(d0, cga);
        d1 = cscale(d1, cga);
        d2 = cscale(d2, cga);
      }
      if (DB) {
        const float2 ph = g.phx[kx];
        d0 = cadd(d0, cmul(ph, DB[eo]));
        d1 = cadd(d1, cmul(ph, DB[eo + TM * A]));
        d2 = cadd(d2, cmul(ph, DB[eo + 2 * TM * A]));
        if constexpr (CG) {
          DB[eo] = cmulc(d0, ph);
          DB[eo + TM * A] = cmulc(d1, ph);
          DB[eo + 2 * TM * A] = cmulc(d2, ph);
        }
      }
      const double e2m = (double)d0.x * d0.x + (double)d0.y * d0.y + (double)d1.x * d1.x + (double)d1.y * d1.y +
                         (double)d2.x * d2.x + (double)d2.y * d2.y;
      acc += (double)coef[m][3] * e2m;
      if (!do_gamma) continue;
      const float k0 = coef[m][0] * g.sx[kx];
      const float kk1 = coef[m][1] * g.cx[kx];
      const float kk2 = coef[m][2] * g.cx[kx];
      const float q2 = k0 * k0 + kk1 * kk1 + kk2 * kk2;
      float2 u0 = make_float2(0.f, 0.f), u1 = u0, u2 = u0;
      if (q2 > 0.f) {
        const float inv = 1.0f / sqrtf(q2);
        const float n0 = k0 * inv, n1 = kk1 * inv, n2 = kk2 * inv;
        const float rhox = (n0 * d0.x + n1 * d1.x + n2 * d2.x) * inv, rhoy = (n0 * d0.y + n1 * d1.y + n2 * d2.y) * inv;
        if constexpr (CG)
          acc2 += (double)coef[m][3] *
                  (e2m / (double)q2 - (double)ccg * ((double)rhox * rhox + (double)rhoy * rhoy));
        const float cA = d.cA * inv * inv, cB = d.cB * inv;
        u0 = make_float2(cA * d0.x + cB * rhox * n0, cA * d0.y + cB * rhoy * n0);
        u1 = make_float2(cA * d1.x + cB * rhox * n1, cA * d1.y + cB * rhoy * n1);
        u2 = make_float2(cA * d2.x + cB * rhox * n2, cA * d2.y + cB * rhoy * n2);
        if (rec) aa_y(coef[m][3], q2, inv, n0, n1, n2, u0, u1, u2, Un[eo], Un[eo + TM * A], Un[eo + 2 * TM * A]);
      } else if (rec) {
        Un[eo] = Un[eo + TM * A] = Un[eo + 2 * TM * A] = make_float2(0.f, 0.f);
      }
      W[p0] = u0;
      W[p1] = u1;
      W[p2] = u2;
    }
    if (do_gamma) {
      __syncthreads();
      for (int it = threadIdx.x; it < A * LH; it += NT) {
        const int line = it / A, k1 = it - line * A;
        if (!valid[line / 3]) continue;
        float2 u[B];
#pragma unroll
        for (int k2 = 0; k2 < B; ++k2) u[k2] = W[L::at(line, k2 + B * k1)];
        dft_reg<B, true>(u);
#pragma unroll
        for (int t = 1; t < B; ++t) u[t] = cmulc(u[t], tw[t * k1]);
#pragma unroll
        for (int t = 0; t < B; ++t) W[L::at(line, t + B * k1)] = u[t];
      }
    }
  }
  const double bs = block_sum<NT>(acc, red);
  if (threadIdx.x == 0) partials[(long long)r * ntiles + tile] = bs;
  if constexpr (CG) {
    const double b2 = block_sum<NT>(acc2, red);
    if (threadIdx.x == 0) aa.cg_part[(long long)r * ntiles + tile] = b2;
  }
  if (!do_gamma) return;
  __syncthreads();

  // stage 3 (inverse): item (t, line): A-point inverse DFTs, g(x = t + B j) into H
  if constexpr (INPL) {
    const int it = threadIdx.x;
    const int t = it / LH, l = it - t * LH;
    float2 v[A];
    if (it < B * LH) {
#pragma unroll
      for (int k1 = 0; k1 < A; ++k1) v[k1] = W[L::at(l, t + B * k1)];
      dft_reg<A, true>(v);
    }
    __syncthreads();
    if (it < B * LH) {
#pragma unroll
      for (int j = 0; j < A; ++j) H[(t + B * j) * LH + l] = v[j];
    }
  } else {
    for (int it = threadIdx.x; it < B * LH; it += NT) {
      const int t = it / LH, l = it - t * LH;
      float2 v[A];
#pragma unroll
      for (int k1 = 0; k1 < A; ++k1) v[k1] = W[L::at(l, t + B * k1)];
      dft_reg<A, true>(v);
#pragma unroll
      for (int j = 0; j < A; ++j) H[(t + B * j) * LH + l] = v[j];
    }
  }
  __syncthreads();

  // stage 4: item (x, m): the six strain components from g(x), g(x - 1)
  for (int it = threadIdx.x; it < N * TM; it += NT) {
    const int x = it / TM, m = it - x * TM;
    if (!valid[m]) continue;
    const float al = 0.5f * coef[m][0], be = 0.5f * coef[m][1], ga = 0.5f * coef[m][2];
    const int xm = (x == 0) ? N - 1 : x - 1;
    const float2 gx0 = H[x * LH + 3 * m], gy0 = H[x * LH + 3 * m + 1], gz0 = H[x * LH + 3 * m + 2];
    const float2 gx1 = H[xm * LH + 3 * m], gy1 = H[xm * LH + 3 * m + 1], gz1 = H[xm * LH + 3 * m + 2];
    // D_j = (g_j(x) - g_j(x-1)) / 2i = ((dy)/2, -(dx)/2); Av_j = (g_j(x) + g_j(x-1)) / 2
    const float2 Dx = make_float2(0.5f * (gx0.y - gx1.y), -0.5f * (gx0.x - gx1.x));
    const float2 Dy = make_float2(0.5f * (gy0.y - gy1.y), -0.5f * (gy0.x - gy1.x));
    const float2 Dz = make_float2(0.5f * (gz0.y - gz1.y), -0.5f * (gz0.x - gz1.x));
    const float2 Ax = make_float2(gx0.x + gx1.x, gx0.y + gx1.y);  // 2 Av
    const float2 Ay = make_float2(gy0.x + gy1.x, gy0.y + gy1.y);
    const float2 Az = make_float2(gz0.x + gz1.x, gz0.y + gz1.y);
    // r_xx = alpha D_x, r_yy = beta Av_y, r_zz = gamma Av_z,
    // r_xy = (alpha D_y + beta Av_x) / 2, r_xz = (alpha D_z + gamma Av_x) / 2,
    // r_yz = (beta Av_z + gamma Av_y) / 2   (al, be, ga hold half the coefficients)
    const float4 o01 = make_float4(2.f * al * Dx.x, 2.f * al * Dx.y, be * Ay.x, be * Ay.y);
    const float4 o23 = make_float4(ga * Az.x, ga * Az.y, al * Dy.x + 0.5f * be * Ax.x, al * Dy.y + 0.5f * be * Ax.y);
    const float4 o45 = make_float4(al * Dz.x + 0.5f * ga * Ax.x, al * Dz.y + 0.5f * ga * Ax.y,
                                   0.5f * (be * Az.x + ga * Ay.x), 0.5f * (be * Az.y + ga * Ay.y));
    float4* o = reinterpret_cast<float4*>(S + x * colmax + 6 * m);
    o[0] = o01;
    o[1] = o23;
    o[2] = o45;
  }
}

}  // namespace lsk
