// Phase-field fixed point of phase_field_solve
// (src/jaxmaterials/solver/lippmann_schwinger.py:465-547), in FP64 on the
// solver's grid and FFT plans:
//
//   A = 1/lc^2 + B,  B = 2 H / (gc lc) (evaluated as (2 H) / (gc lc)),
//   A0 = (min A + max A) / 2,
//   chi_0 = B - (A - A0) d_old,
//   d_k+1 = Re IFFT[ FFT(chi_k) / (A0 + |xi~|^2) ],
//   chi_k+1 = B - (A - A0) d_k+1,
//
// stopped when ||chi_k+1 - chi_k|| / ||chi_k+1|| <= tolerance (FP64 norms).
// The fields are full complex along every axis (the reference uses fftn);
// five passes per iteration: z forward (from the real chi), y forward,
// x forward + division by A0 + |xi~|^2 + x inverse, y inverse, z inverse
// with the chi update and the block partials of the two norms.  A block
// holds up to kPfItems complex numbers of lines in LDS (the LDS Stockham
// engine in FP64, any length up to kPfItems).
#pragma once
#include "ls_kernels.hpp"

namespace lsk {

constexpr int kPfE = 16;
constexpr int kPfNT = 256;
// lines x length per block: ceil(items / R / NT) R <= E for every radix
// R <= 5 of the engine and the direct DFT
constexpr int kPfItems = 3840;

// per-axis FP64 tables of the phase-field symbol: s = sin(pi k / n),
// c = cos(pi k / n) (exact 0 at k = 0 and k = n / 2), g = 2 n / L
struct PfTables {
  const double *sx, *cx, *sy, *cy, *sz, *cz;
  double gx, gy, gz;
};

__device__ __forceinline__ double pf_xixi(const PfTables& t, int kx, int ky, int kz) {
  const double a = t.gx * t.sx[kx] * t.cy[ky] * t.cz[kz];
  const double b = t.gy * t.cx[kx] * t.sy[ky] * t.cz[kz];
  const double c = t.gz * t.cx[kx] * t.cy[ky] * t.sz[kz];
  return a * a + b * b + c * c;
}

// chi = B - (A - A0) d_old, d = d_old
__global__ void __launch_bounds__(kThreads) k_pf_init(long long N, const double* __restrict__ H,
                                                      const double* __restrict__ d_old, double inv_lc2, double gclc,
                                                      double A0, double* __restrict__ chi, double* __restrict__ d) {
  const long long st = (long long)gridDim.x * blockDim.x;
  for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < N; v += st) {
    const double B = 2.0 * H[v] / gclc;
    const double A = inv_lc2 + B;
    chi[v] = B - (A - A0) * d_old[v];
    d[v] = d_old[v];
  }
}

// z lines, forward, from the real chi: L lines of nz per block
__global__ void __launch_bounds__(kPfNT) k_pf_z_fwd(GridDev g, const int* __restrict__ stop,
                                                    const double* __restrict__ chi, double2* __restrict__ C, int L) {
  if (*stop) return;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  double2* buf = reinterpret_cast<double2*>(lds_raw);
  const int nz = g.nz, ls = g.lsz;
  const int nl = g.nx * g.ny;
  const int line0 = blockIdx.x * L;
  const int lines = min(L, nl - line0);
  for (int e = threadIdx.x; e < lines * nz; e += kPfNT) {
    const int l = e / nz, k = e - l * nz;
    buf[lds_idx<true>(l, k, 1, ls)] = make_double2(chi[(long long)(line0 + l) * nz + k], 0.0);
  }
  __syncthreads();
  fft_lines<false, false, true, kPfE, kPfNT, false>(buf, g.pz, g.f_nz, lines, 1, ls);
  for (int e = threadIdx.x; e < lines * nz; e += kPfNT) {
    const int l = e / nz, k = e - l * nz;
    C[(long long)(line0 + l) * nz + k] = buf[lds_idx<true>(l, k, 1, ls)];
  }
}

// z lines, inverse, then d = Re / N, chi_new = B - (A - A0) d and the block
// partials of sum (chi_new - chi)^2 and sum chi_new^2
__global__ void __launch_bounds__(kPfNT) k_pf_z_inv(GridDev g, const int* __restrict__ stop,
                                                    const double2* __restrict__ C, const double* __restrict__ H,
                                                    double inv_lc2, double gclc, double A0, double* __restrict__ chi,
                                                    double* __restrict__ d, double* __restrict__ part, int L) {
  if (*stop) return;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  double2* buf = reinterpret_cast<double2*>(lds_raw);
  __shared__ double red[kPfNT / 64];
  const int nz = g.nz, ls = g.lsz;
  const int nl = g.nx * g.ny;
  const int line0 = blockIdx.x * L;
  const int lines = min(L, nl - line0);
  for (int e = threadIdx.x; e < lines * nz; e += kPfNT) {
    const int l = e / nz, k = e - l * nz;
    buf[lds_idx<true>(l, k, 1, ls)] = C[(long long)(line0 + l) * nz + k];
  }
  __syncthreads();
  fft_lines<true, false, true, kPfE, kPfNT, false>(buf, g.pz, g.f_nz, lines, 1, ls);
  const double invN = 1.0 / (double)g.N;
  double s_diff = 0.0, s_new = 0.0;
  for (int e = threadIdx.x; e < lines * nz; e += kPfNT) {
    const int l = e / nz, k = e - l * nz;
    const long long v = (long long)(line0 + l) * nz + k;
    const double dv = buf[lds_idx<true>(l, k, 1, ls)].x * invN;
    const double B = 2.0 * H[v] / gclc;
    const double A = inv_lc2 + B;
    const double cn = B - (A - A0) * dv;
    const double df = cn - chi[v];
    s_diff += df * df;
    s_new += cn * cn;
    chi[v] = cn;
    d[v] = dv;
  }
  const double a = block_sum<kPfNT>(s_diff, red);
  if (threadIdx.x == 0) part[2 * blockIdx.x] = a;
  const double b = block_sum<kPfNT>(s_new, red);
  if (threadIdx.x == 0) part[2 * blockIdx.x + 1] = b;
}

// y lines of T consecutive z columns at one x
template <bool INV>
__global__ void __launch_bounds__(kPfNT) k_pf_y(GridDev g, const int* __restrict__ stop, double2* __restrict__ C,
                                                FastDiv f_T, int nchunks) {
  if (*stop) return;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  double2* buf = reinterpret_cast<double2*>(lds_raw);
  const int T = (int)f_T.d;
  const int ny = g.ny, nz = g.nz;
  const int x = blockIdx.x / nchunks;
  const int z0 = (blockIdx.x - x * nchunks) * T;
  const int nt = min(T, nz - z0);
  double2* S = C + (long long)x * ny * nz + z0;
  for (int e = threadIdx.x; e < ny * T; e += kPfNT) {
    const int y = fdiv(e, f_T), t = e - y * T;
    buf[e] = t < nt ? S[(long long)y * nz + t] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  fft_lines<INV, true, false, kPfE, kPfNT>(buf, g.py, f_T, T, T, 1);
  for (int e = threadIdx.x; e < ny * T; e += kPfNT) {
    const int y = fdiv(e, f_T), t = e - y * T;
    if (t < nt) S[(long long)y * nz + t] = buf[e];
  }
}

// x lines of T consecutive (y, z) columns: forward, divide by A0 + |xi~|^2,
// inverse
__global__ void __launch_bounds__(kPfNT) k_pf_x(GridDev g, const int* __restrict__ stop, PfTables tb,
                                                double2* __restrict__ C, FastDiv f_T, double A0) {
  if (*stop) return;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  double2* buf = reinterpret_cast<double2*>(lds_raw);
  const int T = (int)f_T.d;
  const int nx = g.nx, nz = g.nz;
  const long long Q = (long long)g.ny * nz;
  const long long q0 = (long long)blockIdx.x * T;
  const int nt = (int)min((long long)T, Q - q0);
  for (int e = threadIdx.x; e < nx * T; e += kPfNT) {
    const int x = fdiv(e, f_T), t = e - x * T;
    buf[e] = t < nt ? C[(long long)x * Q + q0 + t] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  fft_lines<false, true, false, kPfE, kPfNT>(buf, g.px, f_T, T, T, 1);
  for (int e = threadIdx.x; e < nx * T; e += kPfNT) {
    const int x = fdiv(e, f_T), t = e - x * T;
    if (t < nt) {
      const long long q = q0 + t;
      const int ky = (int)(q / nz), kz = (int)(q - (long long)ky * nz);
      const double den = A0 + pf_xixi(tb, x, ky, kz);
      buf[e] = make_double2(buf[e].x / den, buf[e].y / den);
    }
  }
  __syncthreads();
  fft_lines<true, true, false, kPfE, kPfNT>(buf, g.px, f_T, T, T, 1);
  for (int e = threadIdx.x; e < nx * T; e += kPfNT) {
    const int x = fdiv(e, f_T), t = e - x * T;
    if (t < nt) C[(long long)x * Q + q0 + t] = buf[e];
  }
}

// sums of the block partials in a fixed order: out[0..1]
__global__ void __launch_bounds__(kThreads) k_pf_sum(const int* __restrict__ stop, const double* __restrict__ part,
                                                     int nblk, double* __restrict__ out) {
  if (*stop) return;
  __shared__ double red[kThreads / 64];
  for (int c = 0; c < 2; ++c) {
    double a = 0.0;
    for (int i = threadIdx.x; i < nblk; i += kThreads) a += part[2 * i + c];
    const double s = block_sum(a, red);
    if (threadIdx.x == 0) out[c] = s;
  }
}

// Stop test after an iteration, as phase_field_solve's loop condition
// (lippmann_schwinger.py:525-530, 540-545): res = ||chi_new - chi|| / ||chi_new||, or 0
// when ||chi_new|| <= 1e-12; state[0] = stop, state[1] = iterations done.
// Once stop is set the kernels of the remaining iterations of a replayed
// chunk return at once, so the fields and the count are those of the
// one-iteration-at-a-time loop.
__global__ void k_pf_check(const double* __restrict__ sums, double tol, int maxiter, int* __restrict__ state,
                           double* __restrict__ res) {
  if (threadIdx.x != 0 || state[0]) return;
  const double nd = sqrt(sums[0]), nc = sqrt(sums[1]);
  const double r = nc > 1e-12 ? nd / nc : 0.0;
  *res = r;
  state[1] += 1;
  if (!(r > tol) || state[1] >= maxiter) state[0] = 1;
}

// min and max of a FP64 field: per block, then one block over the partials
__global__ void __launch_bounds__(kThreads) k_pf_minmax(const double* __restrict__ f, long long N,
                                                        double* __restrict__ part) {
  __shared__ double rmin[kThreads], rmax[kThreads];
  double lo = INFINITY, hi = -INFINITY;
  const long long st = (long long)gridDim.x * blockDim.x;
  for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < N; v += st) {
    lo = fmin(lo, f[v]);
    hi = fmax(hi, f[v]);
  }
  rmin[threadIdx.x] = lo;
  rmax[threadIdx.x] = hi;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      rmin[threadIdx.x] = fmin(rmin[threadIdx.x], rmin[threadIdx.x + w]);
      rmax[threadIdx.x] = fmax(rmax[threadIdx.x], rmax[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = rmin[0];
    part[2 * blockIdx.x + 1] = rmax[0];
  }
}

}  // namespace lsk
