// HIP kernels of the Lippmann-Schwinger iteration for gfx950.
//
// Formulation.  The strain is held as eps = base + delta.  The FP32 iteration
// transforms only the correction's stress C delta; the divergence of the base
// stress, d_base = xi~ . FFT(C base + tau), is evaluated in FP64 once per base
// and stored.  Because Gamma0 sigma_hat depends on sigma_hat only through the
// divergence xi~ . sigma_hat, the iteration
//     d = d_base + xi~ . FFT32(C delta);  R = ||d|| / ||sigma_hat(0)||;
//     delta += IFFT32(Gamma0(d))
// is the reference iteration eps <- eps - Gamma0 C eps
// (src/jaxmaterials/solver/lippmann_schwinger.py:208-246,
//  cuda/src/lippmann_schwinger.cu:192-216) with the large stress of the base
// kept out of the FP32 FFTs.  FP32 FFT rounding scales with the norm of the
// transformed field; transforming C eps directly leaves a floor near 1e-6 of
// the initial residual at 256^3, transforming C delta does not.  When the
// residual has dropped by a fixed factor the host moves delta into the base
// (k_rebase) and re-evaluates d_base (k_base_z, k_ypass<double2>, k_base_x).
//
// Data layout in HBM (one solver instance, R load cases):
//   lambda, mu : [nmat][N] fp32, C order (x slowest, z fastest)
//   base, delta: [R][6][N] fp32, Voigt (xx,yy,zz,xy,xz,yz), tensor shear
//   tau        : [R][6][N] fp32, optional polarisation (part of the base stress)
//   spec       : [R][nx][ny][P][6] complex fp32, P = nz/2+1 (z half spectrum),
//                the six Voigt components of one mode adjacent (48 B)
//   dbase      : [R][nx][ny*P][3] complex fp32, divergence of the base stress
// One FP32 iteration is four passes over the spectrum:
//   k_zline   : per (x,y) line: C2R along z -> delta update -> Hooke -> R2C along z
//   k_ypass   : forward FFT along y
//   k_xpass   : forward FFT along x -> divergence, residual, Gamma0 -> inverse FFT along x
//   k_ypass   : inverse FFT along y
// followed by k_check (deterministic FP64 residual reduction, stop test).
// With Anderson acceleration the x pass also records the residual's spectral
// representation and its FP64 Gram row, so the weights are known before the
// z pass, which then updates, mixes, applies Hooke and transforms in one go.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "damage.hpp"
#include "fft_lds.hpp"

namespace lsk {

constexpr int kMaxDepth1 = 9;  // Anderson depth <= 8

struct GridDev {
  int nx, ny, nz, P;
  long long N;  // nx*ny*nz
  long long M;  // nx*ny*P
  float gx, gy, gz;  // 2 n_d / L_d
  const float* sx;   // sin(pi k / n) per axis, exact 0 at k = 0
  const float* cx;   // cos(pi k / n) per axis, exact 0 at k = n/2
  const float* sy;
  const float* cy;
  const float* sz;
  const float* cz;
  const float2* phx;  // e^{+i pi k / nx}, k < nx (x pass with three-line transforms)
  AxisPlan px, py, pz;
  FastDiv f_nz, f_P;  // divisions by nz and P
  int lsz;            // padded LDS line stride of the z-line kernels
  // phase-field damage law (null dmg: Hooke's law): per-material damage
  // fields [nmat][N], the stability constant k, and the load cases' bases
  const float* dmg;
  float kdmg;
  const float* base;
};

// Layout of d_base matched to the tiles of the FP32 x pass: [tile][x][m][3]
// with tile = gm / Tm and m = gm % Tm, so each x-pass block reads one
// contiguous range.
// Two orders inside a tile of nx * Tm * 3 complex entries:
//  * LDS-engine x pass (hA = 0): [x][m][3];
//  * three-line x pass (hA = A, nx = A B, x = k1 + A k2): [k2][3][m][k1], so
//    the threads of its stage 2 (consecutive k1) read consecutive entries.
// index() gives component 0, cstride() the step between the components.
struct DbLayout {
  FastDiv f_Tm;        // modes per tile of the FP32 x pass
  FastDiv f_A;         // three-line x pass: A
  int hA;              // 0 or A
  int nx;
  long long per_rhs;  // complex entries per load case
  __device__ __forceinline__ long long index(int x, int gm) const {
    const int t = fdiv(gm, f_Tm);
    const int Tm = (int)f_Tm.d;
    const int m = gm - t * Tm;
    if (hA == 0) return (((long long)t * nx + x) * (long long)Tm + m) * 3;
    const int k2 = fdiv(x, f_A), k1 = x - k2 * hA;
    return (long long)t * nx * Tm * 3 + ((long long)k2 * 3 * Tm + m) * hA + k1;
  }
  __device__ __forceinline__ int cstride() const { return hA == 0 ? 1 : (int)f_Tm.d * hA; }
};

struct RhsDesc {
  int mat;          // index of the (lambda, mu) pair
  float eps_bar[6];
  float cA, cB;     // Gamma0 constants, 1/N of the unnormalised inverse folded in
  float lam0, mu0;  // reference medium
};

struct RhsState {
  double R0;      // normaliser of the relative tolerance
  double R;       // latest residual
  double den;     // ||sigma_hat(0)||_V of the latest check
  double Rbase;   // residual when the current base was formed
  double dcb[6];  // sigma_hat(0) of the base stress (FP64)
  int checks;     // residual checks done
  int active;     // 1 while the load case iterates
  int converged;
  int status;  // 0 ok, 1 non-finite residual
};

// z-line pass modes
enum ZMode : int {
  Z_UPD = 2,    // C2R -> delta += r -> sigma = C delta -> R2C
  Z_SET = 3,    // C2R -> eps = eps_bar - r (end of a projection; no forward transform)
  Z_PROJ = 4,   // eps read, sigma = C0 (eps - eps_bar), R2C  (start of a projection)
  Z_AA = 5,     // C2R -> r; g = delta + r into the newest slot; next delta = sum_s alpha_s g^(s) -> C delta -> R2C
  Z_CG = 6,     // C2R -> r; delta += alpha p; p = r + beta p -> C p -> R2C, block partials of <p, C p>_w
};

// Conjugate gradients on the same equation (LS_CG).  On compatible fields
// with zero mean, A = Gamma0 C is self-adjoint and positive definite in the
// reference-medium product <a, b>_C0 = <a, C0 b>_w (Gamma0 C0 is the
// C0-orthogonal projection onto them), and the Lippmann-Schwinger update
// r = -Gamma0 (C eps) is the residual b - A e of eps = eps_bar + e.  Per load
// case and iteration k:
//   z pass     r_k from the inverse transform; delta += alpha_{k-1} p_{k-1}
//              (the update of the previous step, applied one pass late);
//              p_k = r_k + beta_{k-1} p_{k-1}; sigma = C p_k, R2C; and
//              <p_k, A p_k>_C0 = <p_k, C p_k>_w (pointwise, Gamma0 C0 p = p);
//   k_cg_alpha alpha_k = rr_k / <p_k, C p_k>_w;
//   x pass     d += alpha_k xi~ . (C p_k)^ (d, the divergence of the
//              iterate's stress, kept in the base-divergence buffer: at a
//              (re)start it is the FP64 base divergence); the residual R from
//              d; rr_{k+1} = <r, C0 r>_w from d per mode,
//              sum_k w_k (|d|^2 - c |n . d|^2) / (mu0 |xi~|^2) / N,
//              c = (lambda0 + mu0) / (lambda0 + 2 mu0) (r_hat = -Gamma0
//              sigma_hat is a function of d; the self-conjugate planes need no
//              correction: r_hat of a real stress is Hermitian there);
//              r_hat_{k+1} = -Gamma0 from d;
//   k_cg_beta  beta_k = rr_{k+1} / rr_k (0 at a restart).
// A rebase (the FP64 re-evaluation of the correction formulation) restarts
// the recurrence from the exact divergence of the new base.
struct CgState {
  double alpha;  // the pending update delta += alpha p (not yet applied when applied == 0)
  double beta;
  double rr;     // <r, C0 r>_w of the current residual
  int restart;   // next beta is 0
  int applied;   // the pending update has been applied (by a z pass or k_cg_finish)
};

// Anderson history (lippmann_schwinger.py:177-231, doc/linear_elasticity.tex:335-370).
// Per load case a ring of g^(s) = delta^(s) + r^(s) in real space (the
// mixing operands) and a ring of three complex numbers per mode from which
// the Gram row <r, r^(s)>_w is a plain dot product.  With
// r_hat = sym(xi~ (x) u) (u from the x pass), Parseval and the Hermitian
// weight w_k of the half spectrum,
//   <r_hat, r_hat'>_w = w_k (|xi~|^2 u . conj(u') + (xi~ . u) conj(xi~ . u')) / 2
//                     = Re(y . conj(y')),
//   y = sqrt(w_k / 2) |xi~| (u + (sqrt(2) - 1) (n . u) n),  n = xi~ / |xi~|
// (y = sqrt(w_k/2) Q^{1/2} u with Q = |xi~|^2 I + xi~ xi~^T), so the x pass
// stores y and k_aa_gram streams the ring.  Entries outside the grid's modes
// stay zero.
//
// The correction delta of the current iterate is not stored during the
// iteration: delta^(k) = sum_s alpha^(k-1)_s g^(s) is rebuilt from the ring,
// whose operands are all still there when the z pass of iteration k runs
// (the newest slot keeps the oldest entry until that pass overwrites it).
// Every such sum runs oldest slot first, newest last, with the same FP32
// operations in the z pass that forms delta, in the z pass and k_rebase that
// rebuild it and in k_aa_delta that writes it out, so the rebuilt value is
// bit-identical to the one whose stress was transformed.
struct AAParams {
  float* G;             // [R][m1][6][N]
  float2* U;            // [R][m1][u_per_rhs] complex (y per mode, three per mode)
  long long u_per_rhs;  // complex entries per slot (= DbLayout::per_rhs)
  const double* alpha;  // [R][m1] mixing weights by slot
  const double* alpha_prev;  // [R][m1] weights of the previous mix: delta of the current iterate
  const int* head;      // [R] slot of the newest entry
  const int* filled;    // [R] number of filled slots including the newest
  int m1;               // depth + 1
  int record;           // x pass: store y of the newest residual
  // conjugate gradients: direction p [R][6][N], state per load case, block
  // partials (z pass: <p, C p>_w; x pass: the C0 norm of the residual)
  float* P;
  CgState* cg;
  double* cg_part;
};

// The self-conjugate planes k_z = 0 and (nz even) k_z = nz/2.  There the
// reference's xi~ (k in [0, n)) is odd under the in-plane flip (k_x, k_y) ->
// (-k_x, -k_y), so r_hat is not Hermitian on those planes and the C2R
// transform keeps its Hermitian part only: after the inverse y pass the
// plane holds R(x, y) whose real part is the field and whose imaginary part
// the C2R drops.  The Gram row from y counts |R|^2; the inverse y pass
// records Im R of the two planes in a small ring, and k_aa_gram_planes
// subtracts their contribution, so the Gram row is the real-space one.
struct AAPlanes {
  float* im;            // [R][m1][npl][nx][ny][6] (nullptr: not recording)
  const int* head;      // [R] newest slot
  long long per_slot;   // floats per slot = npl * nx * ny * 6
  int m1;
};

__device__ __forceinline__ void aa_plane_record(const AAPlanes& pl, const GridDev& g, int r, int x, int y, int q,
                                                float im) {
  const int kz = q / 6, c = q - 6 * kz;
  if (kz != 0 && 2 * kz != g.nz) return;
  const int plane = kz == 0 ? 0 : 1;
  pl.im[((long long)r * pl.m1 + pl.head[r]) * pl.per_slot + (((long long)plane * g.nx + x) * g.ny + y) * 6 + c] = im;
}

// y of one mode (see AAParams); q2 = |xi~|^2 > 0, inv = 1 / |xi~|
__device__ __forceinline__ void aa_y(float wk, float q2, float inv, float n0, float n1, float n2, float2 u0,
                                     float2 u1, float2 u2, float2& y0, float2& y1, float2& y2) {
  constexpr float c = 0.41421356237309504880f;  // sqrt(2) - 1
  const float sc = sqrtf(0.5f * wk) * q2 * inv;  // sqrt(w/2) |xi~|
  const float pr = c * (n0 * u0.x + n1 * u1.x + n2 * u2.x), pi = c * (n0 * u0.y + n1 * u1.y + n2 * u2.y);
  y0 = make_float2(sc * (u0.x + pr * n0), sc * (u0.y + pi * n0));
  y1 = make_float2(sc * (u1.x + pr * n1), sc * (u1.y + pi * n1));
  y2 = make_float2(sc * (u2.x + pr * n2), sc * (u2.y + pi * n2));
}

__device__ __forceinline__ float4 ld4(const float2* p) { return *reinterpret_cast<const float4*>(p); }

// Streaming accesses with the non-temporal hint, used by the z passes for the
// data they touch once (spectrum records, strain and ring fields): on MI355X a
// plain grid-stride float4 copy moves 4.9 TB/s, 5.2 TB/s with non-temporal
// loads and stores, and the Anderson z pass 4.83 -> 5.2 TB/s.  Not in the y
// and x passes: their narrow column accesses share cache lines between work
// items, and the hint halved the register y pass (4.7 -> 2.5 TB/s).
typedef float f2v_t __attribute__((ext_vector_type(2)));
typedef float f4v_t __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float ldnt(const float* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ void stnt(float* p, float a) { __builtin_nontemporal_store(a, p); }
__device__ __forceinline__ void stnt(float2* p, float2 a) {
  const f2v_t v = {a.x, a.y};
  __builtin_nontemporal_store(v, reinterpret_cast<f2v_t*>(p));
}
__device__ __forceinline__ float4 ld4nt(const void* p) {
  const f4v_t v = __builtin_nontemporal_load(reinterpret_cast<const f4v_t*>(p));
  return make_float4(v.x, v.y, v.z, v.w);
}
// the rare FP64 passes keep plain accesses
__device__ __forceinline__ void stnt(double2* p, double2 a) { *p = a; }

// Hooke's law, isotropic: sigma_i = 2 mu eps_i + lambda tr(eps) [i<3]
// (src/jaxmaterials/solver/lippmann_schwinger.py:20-33).
template <typename T>
__device__ __forceinline__ void hooke(const T* e, T lam, T mu, T* s) {
  const T tr = e[0] + e[1] + e[2];
  const T m2 = T(2) * mu;
#pragma unroll
  for (int c = 0; c < 3; ++c) s[c] = m2 * e[c] + lam * tr;
#pragma unroll
  for (int c = 3; c < 6; ++c) s[c] = m2 * e[c];
}

// wave64 / block reductions in FP64 with a fixed order (deterministic)
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <int NT = kThreads>
__device__ __forceinline__ double block_sum(double v, double* red) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
    for (int i = 0; i < NT / 64; ++i) s += red[i];
  }
  return s;  // valid in thread 0
}

// Block sums of the first n <= kMaxDepth1 values of v in one pass (the same
// fixed order as block_sum, one pair of barriers for all of them); valid in
// thread 0.  red: kThreads / 64 * kMaxDepth1 doubles.
__device__ __forceinline__ void block_sum_n(double (&v)[kMaxDepth1], int n, double* red) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < kMaxDepth1; ++q)
    if (q < n) v[q] = wave_sum(v[q]);
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < kMaxDepth1; ++q)
      if (q < n) red[w * kMaxDepth1 + q] = v[q];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < kMaxDepth1; ++q) {
      if (q < n) {
        double s = 0.0;
        for (int i = 0; i < kThreads / 64; ++i) s += red[i * kMaxDepth1 + q];
        v[q] = s;
      }
    }
  }
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2.
// Remap the linear workgroup id so that every XCD works through a contiguous
// range of tiles and neighbouring tiles (adjacent HBM segments) share an L2.
__device__ __forceinline__ int xcd_remap(int b, int n) {
  constexpr int X = 8;
  const int per = n / X, rem = n % X;
  const int x = b % X, k = b / X;
  return x * per + min(x, rem) + k;
}

// Split paired forward transforms z_p = a + i b (LDS lines) into the two half
// spectra and store them at S[line][k][2p], S[line][k][2p+1].
template <int NT, typename V>
__device__ __forceinline__ void store_half_pairs(const V* lds, const GridDev& g, int L, V* S, long long line0) {
  using Sc = typename CT<V>::S;
  const int nz = g.nz, P = g.P, ls = g.lsz;
  const int items = L * P * 3;
  for (int e = threadIdx.x; e < items; e += NT) {
    const int e3 = e / 3;
    const int p = e - 3 * e3;
    const int l = fdiv(e3, g.f_P);
    const int k = e3 - l * P;
    const V zk = lds[lds_idx<true>(l * 3 + p, k, 1, ls)];
    const V zm = lds[lds_idx<true>(l * 3 + p, k == 0 ? 0 : nz - k, 1, ls)];
    V* o = S + ((line0 + l) * P + k) * 6 + 2 * p;
    stnt(o, CT<V>::mk(Sc(0.5) * (zk.x + zm.x), Sc(0.5) * (zk.y - zm.y)));
    stnt(o + 1, CT<V>::mk(Sc(0.5) * (zk.y + zm.y), Sc(-0.5) * (zk.x - zm.x)));
  }
}

// ---------------------------------------------------------------------------
// z-line pass.  Block = Lz consecutive (x,y) lines of one load case.
// LDS: Z[l][p][k], p = component pair (0,1),(2,3),(4,5): z_p = s_{2p} + i s_{2p+1}.
template <int MODE, int E, int NT, bool DMG = false>
__global__ void __launch_bounds__(NT) k_zline(GridDev g, const RhsDesc* __restrict__ rhs,
                                              RhsState* __restrict__ st, const float* __restrict__ lam,
                                              const float* __restrict__ mu, float* __restrict__ eps,
                                              float2* __restrict__ spec, int Lz, int ntiles, AAParams aa) {
  extern __shared__ float2 lds[];
  __shared__ double red[NT / 64];
  const int lin = xcd_remap(blockIdx.x, gridDim.x);
  const int r = lin / ntiles;
  const int tile = lin - r * ntiles;
  if (!st[r].active) return;
  const int nz = g.nz, P = g.P;
  const int nlines_xy = g.nx * g.ny;
  const int line0 = tile * Lz;
  const int L = min(Lz, nlines_xy - line0);
  const RhsDesc d = rhs[r];
  const long long N = g.N;
  float2* S = spec + (long long)r * g.M * 6;
  float* E_ = eps + (long long)r * 6 * N;
  const float* LA = lam + (long long)d.mat * N;
  const float* MU = mu + (long long)d.mat * N;
  const int ls = g.lsz;
  const int tid = threadIdx.x;
  const FastDiv f_none = g.f_nz;  // unused by the line-contiguous layout

  [[maybe_unused]] float cga = 0.f, cgb = 0.f;
  [[maybe_unused]] double pcp = 0.0;
  if constexpr (MODE == Z_CG) {
    cga = (float)aa.cg[r].alpha;
    cgb = (float)aa.cg[r].beta;
  }
  if constexpr (MODE == Z_UPD || MODE == Z_SET || MODE == Z_AA || MODE == Z_CG) {
    // gather the half spectra into the paired full-length lines (C2R semantics:
    // imaginary parts of the k = 0 and k = nz/2 coefficients are ignored)
    const int items = L * P * 3;
    for (int e = tid; e < items; e += NT) {
      const int e3 = e / 3;
      const int p = e - 3 * e3;
      const int l = fdiv(e3, g.f_P);
      const int k = e3 - l * P;
      float4 ab = ld4nt(S + ((long long)(line0 + l) * P + k) * 6 + 2 * p);
      float2 A = make_float2(ab.x, ab.y), B = make_float2(ab.z, ab.w);
      const bool self = (k == 0) || (2 * k == nz);
      if (self) {
        A.y = 0.f;
        B.y = 0.f;
      }
      lds[lds_idx<true>(l * 3 + p, k, 1, ls)] = make_float2(A.x - B.y, A.y + B.x);
      if (!self) lds[lds_idx<true>(l * 3 + p, nz - k, 1, ls)] = make_float2(A.x + B.y, B.x - A.y);
    }
    __syncthreads();
    fft_lines<true, false, true, E, NT, false>(lds, g.pz, f_none, 3 * L, 1, ls);
  }

  // pointwise: update / Hooke, write sigma pairs into LDS
  {
    const int items = L * nz;
    for (int e = tid; e < items; e += NT) {
      const int l = fdiv(e, g.f_nz), j = e - l * nz;
      const long long v = (long long)(line0 + l) * nz + j;
      float ev[6];
      if constexpr (MODE == Z_PROJ) {
#pragma unroll
        for (int c = 0; c < 6; ++c) ev[c] = E_[c * N + v];
      } else if constexpr (MODE == Z_AA) {
        // delta^(k) = sum_s alpha'_s g^(s) rebuilt from the ring (previous
        // weights alpha'; the newest slot still holds the oldest entry, a
        // term of that sum when the ring was full); r from the inverse
        // transform; g = delta^(k) + r into the newest slot;
        // delta^(k+1) = sum_s alpha_s g^(s).  Both sums oldest slot first,
        // each slot read once.
        const int m1 = aa.m1;
        const int f = aa.filled[r];
        const int h = aa.head[r];
        const long long slot_stride = 6 * N;
        float* Gb = aa.G + (long long)r * m1 * slot_stride;
        const double* al = aa.alpha + (long long)r * m1;
        const double* ap = aa.alpha_prev + (long long)r * m1;
        float* Gh = Gb + h * slot_stride;
        float dk[6], dn[6];
        const double aph = ap[h];
        if (aph != 0.0) {
          const float a = (float)aph;
#pragma unroll
          for (int c = 0; c < 6; ++c) dk[c] = a * ldnt(&Gh[c * N + v]);
        } else {
#pragma unroll
          for (int c = 0; c < 6; ++c) dk[c] = 0.f;
        }
#pragma unroll
        for (int c = 0; c < 6; ++c) dn[c] = 0.f;
        for (int s2 = f - 1; s2 >= 1; --s2) {
          int sl = h - s2;
          if (sl < 0) sl += m1;
          const float a = (float)al[sl], b = (float)ap[sl];
          const float* Gs = Gb + sl * slot_stride;
#pragma unroll
          for (int c = 0; c < 6; ++c) {
            const float gv = ldnt(&Gs[c * N + v]);
            dn[c] = fmaf(a, gv, dn[c]);
            dk[c] = fmaf(b, gv, dk[c]);
          }
        }
        float gn[6];
#pragma unroll
        for (int p = 0; p < 3; ++p) {
          const float2 z = lds[lds_idx<true>(l * 3 + p, j, 1, ls)];
          gn[2 * p] = dk[2 * p] + z.x;
          gn[2 * p + 1] = dk[2 * p + 1] + z.y;
        }
        const float a0 = (float)al[h];
#pragma unroll
        for (int c = 0; c < 6; ++c) {
          stnt(&Gh[c * N + v], gn[c]);
          ev[c] = fmaf(a0, gn[c], dn[c]);
        }
      } else if constexpr (MODE == Z_CG) {
        // r from the inverse transform; delta += alpha p (pending update);
        // p = r + beta p, whose stress is transformed
        float* Pr = aa.P + (long long)r * 6 * N;
#pragma unroll
        for (int p = 0; p < 3; ++p) {
          const float2 z = lds[lds_idx<true>(l * 3 + p, j, 1, ls)];
          ev[2 * p] = z.x;
          ev[2 * p + 1] = z.y;
        }
#pragma unroll
        for (int c = 0; c < 6; ++c) {
          const float po = ldnt(&Pr[c * N + v]);
          stnt(&E_[c * N + v], fmaf(cga, po, ldnt(&E_[c * N + v])));
          ev[c] = fmaf(cgb, po, ev[c]);
          stnt(&Pr[c * N + v], ev[c]);
        }
      } else {
        float rr[6];
#pragma unroll
        for (int p = 0; p < 3; ++p) {
          const float2 z = lds[lds_idx<true>(l * 3 + p, j, 1, ls)];
          rr[2 * p] = z.x;
          rr[2 * p + 1] = z.y;
        }
        if constexpr (MODE == Z_UPD) {
#pragma unroll
          for (int c = 0; c < 6; ++c) ev[c] = ldnt(&E_[c * N + v]) + rr[c];
        } else {
#pragma unroll
          for (int c = 0; c < 6; ++c) ev[c] = d.eps_bar[c] - rr[c];
        }
#pragma unroll
        for (int c = 0; c < 6; ++c) stnt(&E_[c * N + v], ev[c]);
      }
      if constexpr (MODE != Z_SET) {
        float s[6];
        if constexpr (MODE == Z_PROJ) {
          float de[6];
#pragma unroll
          for (int c = 0; c < 6; ++c) de[c] = ev[c] - d.eps_bar[c];
          hooke(de, d.lam0, d.mu0, s);
        } else {
          hooke(ev, LA[v], MU[v], s);
          if constexpr (DMG) {
            // damaged law: sigma(base + delta) - sigma(base)
            //   = C delta + (g - 1) (sigma+(base + delta) - sigma+(base)), in FP64
            const float* B_ = g.base + (long long)r * 6 * N;
            double b[6], t[6], sp1[6], sp0[6];
#pragma unroll
            for (int c = 0; c < 6; ++c) {
              b[c] = B_[c * N + v];
              t[c] = b[c] + (double)ev[c];
            }
            const double lv = LA[v], mv = MU[v];
            sigma_tension(t, lv, mv, sp1);
            sigma_tension(b, lv, mv, sp0);
            const double gm1 = degradation(g.dmg[(long long)d.mat * N + v], g.kdmg) - 1.0;
#pragma unroll
            for (int c = 0; c < 6; ++c) s[c] += (float)(gm1 * (sp1[c] - sp0[c]));
          }
          if constexpr (MODE == Z_CG) {
            pcp += (double)ev[0] * s[0] + (double)ev[1] * s[1] + (double)ev[2] * s[2] +
                   2.0 * ((double)ev[3] * s[3] + (double)ev[4] * s[4] + (double)ev[5] * s[5]);
          }
        }
#pragma unroll
        for (int p = 0; p < 3; ++p) lds[lds_idx<true>(l * 3 + p, j, 1, ls)] = make_float2(s[2 * p], s[2 * p + 1]);
      }
    }
  }
  if constexpr (MODE == Z_CG) {
    const double bs = block_sum<NT>(pcp, red);
    if (tid == 0) {
      aa.cg_part[(long long)r * ntiles + tile] = bs;
      if (tile == 0) aa.cg[r].applied = 1;
    }
  }
  if constexpr (MODE == Z_SET) {
    return;
  } else {
    __syncthreads();
    fft_lines<false, false, true, E, NT, false>(lds, g.pz, f_none, 3 * L, 1, ls);
    __syncthreads();
    store_half_pairs<NT>(lds, g, L, S, (long long)line0);
  }
}

// ---------------------------------------------------------------------------
// y pass: FFT along y of T consecutive (kz, component) columns at fixed x.
// rsel >= 0: the FP64 base evaluation (spec holds that one load case only).
template <bool INV, int E, int NT, typename V>
__global__ void __launch_bounds__(NT) k_ypass(GridDev g, const RhsState* __restrict__ st, V* __restrict__ spec,
                                              FastDiv f_T, int ntiles, int rsel, AAPlanes pl) {
  const int T = (int)f_T.d;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  V* lds = reinterpret_cast<V*>(lds_raw);
  const int lin = xcd_remap(blockIdx.x, gridDim.x);
  const int per_rhs = g.nx * ntiles;
  const int r = rsel >= 0 ? rsel : lin / per_rhs;
  const int tile = rsel >= 0 ? lin : lin - r * per_rhs;
  if (rsel < 0 && !st[r].active) return;
  const int ny = g.ny;
  const int Q = 6 * g.P;  // columns per (x, y)
  const int x = tile / ntiles;
  const int q0 = (tile - x * ntiles) * T;
  V* S = spec + (rsel >= 0 ? 0LL : (long long)r * g.M * 6) + (long long)x * ny * Q;
  const int tid = threadIdx.x;
  const int items = ny * T;
  for (int e = tid; e < items; e += NT) {
    const int y = fdiv(e, f_T), t = e - y * T;
    const int q = q0 + t;
    lds[e] = (q < Q) ? S[(long long)y * Q + q] : CT<V>::mk(0, 0);
  }
  __syncthreads();
  fft_lines<INV, true, false, E, NT>(lds, g.py, f_T, T, T, 1);
  for (int e = tid; e < items; e += NT) {
    const int y = fdiv(e, f_T), t = e - y * T;
    const int q = q0 + t;
    if (q < Q) {
      S[(long long)y * Q + q] = lds[e];
      if constexpr (INV) {
        if (pl.im) aa_plane_record(pl, g, r, x, y, q, (float)lds[e].y);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// xi~(k) from the per-axis tables (fourier.py:9-58 closed form)
__device__ __forceinline__ void xi_tilde(const GridDev& g, int kx, int ky, int kz, float& k0, float& k1,
                                         float& k2) {
  const float sxk = g.sx[kx], cxk = g.cx[kx];
  const float syk = g.sy[ky], cyk = g.cy[ky];
  const float szk = g.sz[kz], czk = g.cz[kz];
  k0 = g.gx * sxk * cyk * czk;
  k1 = g.gy * cxk * syk * czk;
  k2 = g.gz * cxk * cyk * szk;
}

// ---------------------------------------------------------------------------
// x pass: forward FFT along x -> divergence d = d_base + sigma_hat . xi~,
// residual (FP64 partial per block), DC capture, Gamma0(d) -> inverse FFT
// along x.  Block = Tm consecutive (ky,kz) modes of one load case, all six
// components, full x extent.
// Residual (lippmann_schwinger.py:118-143, fourier.cu:140-265):
//   sum_k w_kz |d(k)|^2, w = 1 for kz = 0 or kz = nz/2, else 2.
// Gamma0 (fourier.py:115-186, doc/linear_elasticity.tex:137-226) in terms of
// the divergence: n = xi~/|xi~|, v = d/|xi~| (= sigma_hat n), rho = n . v,
//   r_hat = cA (n0 v0, n1 v1, n2 v2, (n0 v1 + n1 v0)/2, (n0 v2 + n2 v0)/2,
//               (n1 v2 + n2 v1)/2) + cB rho (n0^2, n1^2, n2^2, n0 n1, n0 n2, n1 n2)
// REC: 0 plain; 1 Anderson (record y of the residual); 2 conjugate gradients
// (d = alpha xi~ . sigma_hat + d, written back, and the C0 norm of the residual).
template <int E, int NT, int REC>
__global__ void __launch_bounds__(NT) k_xpass(GridDev g, const RhsDesc* __restrict__ rhs,
                                              const RhsState* __restrict__ st, float2* __restrict__ spec,
                                              float2* __restrict__ dbase, double* __restrict__ partials,
                                              float* __restrict__ dc, FastDiv f_Tm, FastDiv f_cols, int ntiles,
                                              DbLayout dl, int do_gamma, AAParams aa) {
  extern __shared__ float2 lds[];
  __shared__ double red[NT / 64];
  const bool rec = REC == 1 && do_gamma;
  constexpr bool CG = REC == 2;
  const int Tm = (int)f_Tm.d;
  const int lin = xcd_remap(blockIdx.x, gridDim.x);
  const int r = lin / ntiles;
  const int tile = lin - r * ntiles;
  if (!st[r].active) return;
  const int nx = g.nx, P = g.P;
  const long long nm = (long long)g.ny * P;  // modes per x plane
  const int m0 = tile * Tm;
  const int cols = 6 * Tm;
  float2* S = spec + (long long)r * g.M * 6;
  float2* DB = dbase ? dbase + (long long)r * dl.per_rhs + (long long)tile * nx * Tm * 3 : nullptr;
  const RhsDesc d = rhs[r];
  const int tid = threadIdx.x;
  const long long colmax = nm * 6;
  [[maybe_unused]] const float cga = CG ? (float)aa.cg[r].alpha : 1.f;
  [[maybe_unused]] const float ccg = CG ? (d.lam0 + d.mu0) / (d.lam0 + 2.f * d.mu0) : 0.f;
  [[maybe_unused]] double acc2 = 0.0;
  const int items = nx * cols;
  for (int e = tid; e < items; e += NT) {
    const int x = fdiv(e, f_cols), t = e - x * cols;
    const long long q = (long long)m0 * 6 + t;
    lds[e] = (q < colmax) ? S[(long long)x * colmax + q] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  fft_lines<false, true, false, E, NT>(lds, g.px, f_cols, cols, cols, 1);

  double acc = 0.0;
  const int nmodes = nx * Tm;
  for (int e = tid; e < nmodes; e += NT) {
    const int x = fdiv(e, f_Tm), m = e - x * Tm;
    const int gm = m0 + m;
    if (gm >= nm) continue;
    const int ky = fdiv(gm, g.f_P), kz = gm - ky * P;
    float2* t = lds + x * cols + 6 * m;
    float2 s[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) s[c] = t[c];
    if (x == 0 && gm == 0) {
#pragma unroll
      for (int c = 0; c < 6; ++c) dc[r * 6 + c] = s[c].x;
    }
    float k0, k1, k2;
    xi_tilde(g, x, ky, kz, k0, k1, k2);
    // d = d_base + sigma_hat . xi~
    float2 d0, d1, d2;
    d0.x = k0 * s[0].x + k1 * s[3].x + k2 * s[4].x;
    d0.y = k0 * s[0].y + k1 * s[3].y + k2 * s[4].y;
    d1.x = k0 * s[3].x + k1 * s[1].x + k2 * s[5].x;
    d1.y = k0 * s[3].y + k1 * s[1].y + k2 * s[5].y;
    d2.x = k0 * s[4].x + k1 * s[5].x + k2 * s[2].x;
    d2.y = k0 * s[4].y + k1 * s[5].y + k2 * s[2].y;
    if constexpr (CG) {
      d0 = cscale(d0, cga);
      d1 = cscale(d1, cga);
      d2 = cscale(d2, cga);
    }
    if (DB) {
      float2* b = DB + (x * Tm + m) * 3;
      const float2 b0 = b[0], b1 = b[1], b2 = b[2];
      d0.x += b0.x;
      d0.y += b0.y;
      d1.x += b1.x;
      d1.y += b1.y;
      d2.x += b2.x;
      d2.y += b2.y;
      if constexpr (CG) {
        b[0] = d0;
        b[1] = d1;
        b[2] = d2;
      }
    }
    const double wk = (kz == 0 || 2 * kz == g.nz) ? 1.0 : 2.0;
    const double e2m = (double)d0.x * d0.x + (double)d0.y * d0.y + (double)d1.x * d1.x + (double)d1.y * d1.y +
                       (double)d2.x * d2.x + (double)d2.y * d2.y;
    acc += wk * e2m;
    if (do_gamma) {
      const float q2 = k0 * k0 + k1 * k1 + k2 * k2;
      float2 o[6];
      if (q2 > 0.f) {
        const float inv = 1.0f / sqrtf(q2);
        const float n0 = k0 * inv, n1 = k1 * inv, n2 = k2 * inv;
        const float2 v0 = cscale(d0, inv), v1 = cscale(d1, inv), v2 = cscale(d2, inv);
        const float2 rho = make_float2(n0 * v0.x + n1 * v1.x + n2 * v2.x, n0 * v0.y + n1 * v1.y + n2 * v2.y);
        if constexpr (CG)
          acc2 += wk * (e2m / (double)q2 - (double)ccg * ((double)rho.x * rho.x + (double)rho.y * rho.y));
        const float cA = d.cA, cB = d.cB;
        const float h = 0.5f * cA;
        o[0] = make_float2(cA * n0 * v0.x + cB * rho.x * n0 * n0, cA * n0 * v0.y + cB * rho.y * n0 * n0);
        o[1] = make_float2(cA * n1 * v1.x + cB * rho.x * n1 * n1, cA * n1 * v1.y + cB * rho.y * n1 * n1);
        o[2] = make_float2(cA * n2 * v2.x + cB * rho.x * n2 * n2, cA * n2 * v2.y + cB * rho.y * n2 * n2);
        o[3] = make_float2(h * (n0 * v1.x + n1 * v0.x) + cB * rho.x * n0 * n1,
                           h * (n0 * v1.y + n1 * v0.y) + cB * rho.y * n0 * n1);
        o[4] = make_float2(h * (n0 * v2.x + n2 * v0.x) + cB * rho.x * n0 * n2,
                           h * (n0 * v2.y + n2 * v0.y) + cB * rho.y * n0 * n2);
        o[5] = make_float2(h * (n1 * v2.x + n2 * v1.x) + cB * rho.x * n1 * n2,
                           h * (n1 * v2.y + n2 * v1.y) + cB * rho.y * n1 * n2);
        if (rec) {
          // u = (cA v + cB rho n) / |xi~|, recorded as y of e^{+i pi kx/nx} u
          const float2 ph = g.phx[x];
          const float2 u0 = cmul(ph, make_float2((cA * v0.x + cB * rho.x * n0) * inv, (cA * v0.y + cB * rho.y * n0) * inv));
          const float2 u1 = cmul(ph, make_float2((cA * v1.x + cB * rho.x * n1) * inv, (cA * v1.y + cB * rho.y * n1) * inv));
          const float2 u2 = cmul(ph, make_float2((cA * v2.x + cB * rho.x * n2) * inv, (cA * v2.y + cB * rho.y * n2) * inv));
          float2* Un = aa.U + ((long long)r * aa.m1 + aa.head[r]) * aa.u_per_rhs + dl.index(x, gm);
          aa_y((float)wk, q2, inv, n0, n1, n2, u0, u1, u2, Un[0], Un[1], Un[2]);
        }
      } else {
#pragma unroll
        for (int c = 0; c < 6; ++c) o[c] = make_float2(0.f, 0.f);
        if (rec) {
          const long long ui = dl.index(x, gm);
          float2* Un = aa.U + ((long long)r * aa.m1 + aa.head[r]) * aa.u_per_rhs + ui;
          Un[0] = Un[1] = Un[2] = make_float2(0.f, 0.f);
        }
      }
#pragma unroll
      for (int c = 0; c < 6; ++c) t[c] = o[c];
    }
  }
  const double bs = block_sum<NT>(acc, red);
  if (tid == 0) partials[(long long)r * ntiles + tile] = bs;
  if constexpr (CG) {
    const double b2 = block_sum<NT>(acc2, red);
    if (tid == 0) aa.cg_part[(long long)r * ntiles + tile] = b2;
  }
  __syncthreads();
  if (!do_gamma) return;
  fft_lines<true, true, false, E, NT>(lds, g.px, f_cols, cols, cols, 1);
  for (int e = tid; e < items; e += NT) {
    const int x = fdiv(e, f_cols), t = e - x * cols;
    const long long q = (long long)m0 * 6 + t;
    if (q < colmax) S[(long long)x * colmax + q] = lds[e];
  }
}

// ---------------------------------------------------------------------------
// FP64 base evaluation of one load case r (spec64 holds that load case only).
// z lines: sigma = C base + tau in FP64, R2C along z in FP64.
// LM: the uniform base eps_bar of every load case of one material instead:
// lambda and mu of load case r's material in the slots of sigma_0 and
// sigma_1 (the other four zero), since then
// sigma_hat = tr(eps_bar) lambda_hat I + 2 mu_hat eps_bar for each of them
// (k_base_x<LM> forms their divergences).
template <int E, int NT, bool LM = false>
__global__ void __launch_bounds__(NT) k_base_z(GridDev g, const RhsDesc* __restrict__ rhs, int r,
                                               const float* __restrict__ lam, const float* __restrict__ mu,
                                               const float* __restrict__ base, const float* __restrict__ tau,
                                               double2* __restrict__ spec64, int Lz) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  double2* lds = reinterpret_cast<double2*>(lds_raw);
  const int nz = g.nz, P = g.P;
  const int nlines_xy = g.nx * g.ny;
  const int line0 = xcd_remap(blockIdx.x, gridDim.x) * Lz;
  const int L = min(Lz, nlines_xy - line0);
  const RhsDesc d = rhs[r];
  const long long N = g.N;
  const float* B_ = base + (long long)r * 6 * N;
  const float* T_ = tau ? tau + (long long)r * 6 * N : nullptr;
  const float* LA = lam + (long long)d.mat * N;
  const float* MU = mu + (long long)d.mat * N;
  const int ls = g.lsz;
  const int items = L * nz;
  for (int e = threadIdx.x; e < items; e += NT) {
    const int l = fdiv(e, g.f_nz), j = e - l * nz;
    const long long v = (long long)(line0 + l) * nz + j;
    double ev[6], s[6];
    if constexpr (LM) {
      s[0] = LA[v];
      s[1] = MU[v];
#pragma unroll
      for (int c = 2; c < 6; ++c) s[c] = 0.0;
#pragma unroll
      for (int p = 0; p < 3; ++p) lds[lds_idx<true>(l * 3 + p, j, 1, ls)] = make_double2(s[2 * p], s[2 * p + 1]);
      continue;
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) ev[c] = B_[c * N + v];
    hooke(ev, (double)LA[v], (double)MU[v], s);
    if (g.dmg) {
      double sp[6];
      sigma_tension(ev, (double)LA[v], (double)MU[v], sp);
      const double gm1 = degradation(g.dmg[(long long)d.mat * N + v], g.kdmg) - 1.0;
#pragma unroll
      for (int c = 0; c < 6; ++c) s[c] += gm1 * sp[c];
    }
    if (T_) {
#pragma unroll
      for (int c = 0; c < 6; ++c) s[c] += T_[c * N + v];
    }
#pragma unroll
    for (int p = 0; p < 3; ++p) lds[lds_idx<true>(l * 3 + p, j, 1, ls)] = make_double2(s[2 * p], s[2 * p + 1]);
  }
  __syncthreads();
  fft_lines<false, false, true, E, NT, false>(lds, g.pz, g.f_nz, 3 * L, 1, ls);
  __syncthreads();
  store_half_pairs<NT>(lds, g, L, spec64, (long long)line0);
}

// x lines: forward FFT along x in FP64, d_base = sigma_hat . xi~ (the FP32
// tables of the iteration, promoted, so both use the same xi~), base DC.
// LM (after k_base_z<LM>): spec64 holds lambda_hat, mu_hat; d_base and the
// base DC of the load cases r .. r + nr - 1 with sigma_hat =
// tr(eps_bar) lambda_hat I + 2 mu_hat eps_bar.
template <int E, int NT, bool LM = false>
__global__ void __launch_bounds__(NT) k_base_x(GridDev g, int r, const double2* __restrict__ spec64,
                                               float2* __restrict__ dbase, RhsState* __restrict__ st, FastDiv f_Tm,
                                               FastDiv f_cols, DbLayout dl, const RhsDesc* __restrict__ rhs = nullptr,
                                               int nr = 1) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const int Tm = (int)f_Tm.d;
  double2* lds = reinterpret_cast<double2*>(lds_raw);
  const int nx = g.nx, P = g.P;
  const long long nm = (long long)g.ny * P;
  const int m0 = xcd_remap(blockIdx.x, gridDim.x) * Tm;
  const int cols = 6 * Tm;
  const long long colmax = nm * 6;
  const int items = nx * cols;
  for (int e = threadIdx.x; e < items; e += NT) {
    const int x = fdiv(e, f_cols), t = e - x * cols;
    const long long q = (long long)m0 * 6 + t;
    lds[e] = (q < colmax) ? spec64[(long long)x * colmax + q] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  fft_lines<false, true, false, E, NT>(lds, g.px, f_cols, cols, cols, 1);
  const int nmodes = nx * Tm;
  if constexpr (LM) {
    for (int e = threadIdx.x; e < nmodes; e += NT) {
      const int x = fdiv(e, f_Tm), m = e - x * Tm;
      const int gm = m0 + m;
      if (gm >= nm) continue;
      const int ky = fdiv(gm, g.f_P), kz = gm - ky * P;
      const double2 la = lds[x * cols + 6 * m], mu = lds[x * cols + 6 * m + 1];
      float f0, f1, f2;
      xi_tilde(g, x, ky, kz, f0, f1, f2);
      const double k[3] = {f0, f1, f2};
      const int cs = dl.cstride();
      for (int q = r; q < r + nr; ++q) {
        double eb[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) eb[c] = rhs[q].eps_bar[c];
        const double tr = eb[0] + eb[1] + eb[2];
        // sigma_hat_c = (c < 3 ? tr lambda_hat : 0) + 2 mu_hat eps_bar_c
        double2 sg[6];
#pragma unroll
        for (int c = 0; c < 6; ++c)
          sg[c] = make_double2((c < 3 ? tr * la.x : 0.0) + 2.0 * eb[c] * mu.x,
                               (c < 3 ? tr * la.y : 0.0) + 2.0 * eb[c] * mu.y);
        if (x == 0 && gm == 0) {
#pragma unroll
          for (int c = 0; c < 6; ++c) st[q].dcb[c] = sg[c].x;
        }
        float2* o = dbase + (long long)q * dl.per_rhs + dl.index(x, gm);
        o[0] = make_float2((float)(k[0] * sg[0].x + k[1] * sg[3].x + k[2] * sg[4].x),
                           (float)(k[0] * sg[0].y + k[1] * sg[3].y + k[2] * sg[4].y));
        o[cs] = make_float2((float)(k[0] * sg[3].x + k[1] * sg[1].x + k[2] * sg[5].x),
                            (float)(k[0] * sg[3].y + k[1] * sg[1].y + k[2] * sg[5].y));
        o[2 * cs] = make_float2((float)(k[0] * sg[4].x + k[1] * sg[5].x + k[2] * sg[2].x),
                                (float)(k[0] * sg[4].y + k[1] * sg[5].y + k[2] * sg[2].y));
      }
    }
    return;
  }
  float2* DB = dbase + (long long)r * dl.per_rhs;
  for (int e = threadIdx.x; e < nmodes; e += NT) {
    const int x = fdiv(e, f_Tm), m = e - x * Tm;
    const int gm = m0 + m;
    if (gm >= nm) continue;
    const int ky = fdiv(gm, g.f_P), kz = gm - ky * P;
    const double2* s = lds + x * cols + 6 * m;
    if (x == 0 && gm == 0) {
#pragma unroll
      for (int c = 0; c < 6; ++c) st[r].dcb[c] = s[c].x;
    }
    float f0, f1, f2;
    xi_tilde(g, x, ky, kz, f0, f1, f2);
    const double k0 = f0, k1 = f1, k2 = f2;
    float2* o = DB + dl.index(x, gm);
    const int cs = dl.cstride();
    o[0] = make_float2((float)(k0 * s[0].x + k1 * s[3].x + k2 * s[4].x),
                       (float)(k0 * s[0].y + k1 * s[3].y + k2 * s[4].y));
    o[cs] = make_float2((float)(k0 * s[3].x + k1 * s[1].x + k2 * s[5].x),
                        (float)(k0 * s[3].y + k1 * s[1].y + k2 * s[5].y));
    o[2 * cs] = make_float2((float)(k0 * s[4].x + k1 * s[5].x + k2 * s[2].x),
                            (float)(k0 * s[4].y + k1 * s[5].y + k2 * s[2].y));
  }
}

// ---------------------------------------------------------------------------
// Stop test per load case: R = sqrt(sum partials) / ||sigma_hat(0)||_V
// (lippmann_schwinger.py:118-143; stop rule of lippmann_schwinger.cu:205:
// R < max(rtol R0, atol), R0 = residual of the first check unless preset).
// mode 0: a new check; 1: record R only (no stop test); 2: re-evaluation of
// the current check after a rebase (not counted as a new check).
// absolute != 0: R = ||d|| without the ||sigma_hat(0)|| normalisation (the
// adjoint solve, whose stress can have zero mean).
__global__ void __launch_bounds__(kThreads) k_check(RhsState* __restrict__ st, const double* __restrict__ partials,
                                                    const float* __restrict__ dc, int nparts, double rtol,
                                                    double atol, int mode, int absolute) {
  __shared__ double red[kThreads / 64];
  const int r = blockIdx.x;
  if (!st[r].active) return;
  const double* p = partials + (long long)r * nparts;
  double a = 0.0;
  for (int i = threadIdx.x; i < nparts; i += kThreads) a += p[i];
  const double s = block_sum(a, red);
  if (threadIdx.x == 0) {
    RhsState q = st[r];
    const float* z = dc + r * 6;
    double z6[6];
    for (int c = 0; c < 6; ++c) z6[c] = q.dcb[c] + (double)z[c];
    const double den = sqrt(z6[0] * z6[0] + z6[1] * z6[1] + z6[2] * z6[2] +
                            2.0 * (z6[3] * z6[3] + z6[4] * z6[4] + z6[5] * z6[5]));
    const double R = absolute ? sqrt(s) : sqrt(s) / den;
    q.R = R;
    q.den = den;
    if (!(R == R) || R > 1e300) {
      q.status = 1;
      q.active = 0;
    } else if (mode != 1) {
      if (q.checks == 0 && mode == 0) {
        if (!(q.R0 > 0.0)) q.R0 = R;
        q.Rbase = R;
      }
      if (mode == 2) q.Rbase = R;
      const double thr = fmax(rtol * q.R0, atol);
      if (R < thr) {
        q.converged = 1;
        q.active = 0;
      }
    }
    if (mode != 2) q.checks += 1;
    st[r] = q;
  }
}

// ---------------------------------------------------------------------------
// Conjugate-gradient scalars (see CgState).  One block per load case.
// alpha = rr / <p, C p>_w from the z pass partials; a non-positive or
// non-finite <p, C p> (an indefinite medium or a NaN) stops the load case.
__global__ void __launch_bounds__(kThreads) k_cg_alpha(RhsState* __restrict__ st, CgState* __restrict__ cg,
                                                       const double* __restrict__ part, int nparts) {
  __shared__ double red[kThreads / 64];
  const int r = blockIdx.x;
  if (!st[r].active) return;
  const double* p = part + (long long)r * nparts;
  double a = 0.0;
  for (int i = threadIdx.x; i < nparts; i += kThreads) a += p[i];
  const double pcp = block_sum(a, red);
  if (threadIdx.x == 0) {
    CgState c = cg[r];
    if (pcp > 0.0 && pcp < 1e300) {
      c.alpha = c.rr / pcp;
    } else {
      c.alpha = 0.0;
      if (pcp != 0.0) {  // a zero direction (exact solution) is no error
        st[r].status = 1;
        st[r].active = 0;
      }
    }
    c.applied = 0;
    cg[r] = c;
  }
}

// beta = rr_new / rr from the x pass partials (rr_new = sum / (N mu0)); 0 at a restart.
__global__ void __launch_bounds__(kThreads) k_cg_beta(const RhsState* __restrict__ st,
                                                      const RhsDesc* __restrict__ rhs, CgState* __restrict__ cg,
                                                      const double* __restrict__ part, int nparts, double invN) {
  __shared__ double red[kThreads / 64];
  const int r = blockIdx.x;
  if (!st[r].active) return;
  const double* p = part + (long long)r * nparts;
  double a = 0.0;
  for (int i = threadIdx.x; i < nparts; i += kThreads) a += p[i];
  const double s = block_sum(a, red);
  if (threadIdx.x == 0) {
    CgState c = cg[r];
    const double rr = s * invN / (double)rhs[r].mu0;
    c.beta = (c.restart || !(c.rr > 0.0)) ? 0.0 : rr / c.rr;
    c.restart = 0;
    c.rr = rr;
    cg[r] = c;
  }
}

// The pending update delta += alpha p of every load case whose last step was
// an x pass (active or not), with the z pass's FP32 operation.
__global__ void __launch_bounds__(kThreads) k_cg_finish(const CgState* __restrict__ cg, long long N,
                                                        const float* __restrict__ P, float* __restrict__ delta) {
  const int r = blockIdx.y;
  const CgState c = cg[r];
  if (c.applied || c.alpha == 0.0) return;
  const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (i >= 6 * N) return;
  const long long off = (long long)r * 6 * N + i;
  delta[off] = fmaf((float)c.alpha, P[off], delta[off]);
}

// (Re)start of the recurrence for the selected load cases (sel null: all):
// no pending update, the next beta 0.
__global__ void k_cg_reset(CgState* __restrict__ cg, const int* __restrict__ sel, int nrhs) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrhs || (sel && !sel[r])) return;
  CgState c = cg[r];
  c.alpha = 0.0;
  c.beta = 0.0;
  c.restart = 1;
  c.applied = 1;
  cg[r] = c;
}

// ---------------------------------------------------------------------------
// delta of the current iterate from the Anderson ring, sum_s c_s g^(s) with
// c = alpha' (weights of the previous mix) or, with newest_last, c = alpha
// (the mix the last z pass applied: its newest operand sits in slot h).
// aa_ring_load reads each filled slot once (gv[a] = slot h - a, all loads
// issued before the sum); aa_ring_sum adds the terms with c_s != 0 in the
// order of the z pass (oldest first).
__device__ __forceinline__ int aa_slot(int h, int a, int m1) { return h - a < 0 ? h - a + m1 : h - a; }

__device__ __forceinline__ void aa_ring_load(const float* Gb, long long ss, long long i, int h, int f, int m1,
                                             float (&gv)[kMaxDepth1]) {
#pragma unroll
  for (int a = 0; a < kMaxDepth1; ++a) gv[a] = a < f ? Gb[(long long)aa_slot(h, a, m1) * ss + i] : 0.f;
}

__device__ __forceinline__ float aa_ring_sum(const float (&gv)[kMaxDepth1], const double* c, int h, int f, int m1,
                                             bool newest_last) {
  float acc = 0.f;
  if (!newest_last && f > 0 && c[h] != 0.0) acc = (float)c[h] * gv[0];
#pragma unroll
  for (int a = kMaxDepth1 - 1; a >= 1; --a) {
    if (a < f) {
      const double ca = c[aa_slot(h, a, m1)];
      if (ca != 0.0) acc = fmaf((float)ca, gv[a], acc);
    }
  }
  if (newest_last && f > 0 && c[h] != 0.0) acc = fmaf((float)c[h], gv[0], acc);
  return acc;
}

// Move the correction into the base: base += delta, delta = 0.  Plain
// iteration: delta from its buffer.  Anderson (G given): delta rebuilt from
// the ring with the previous weights, and the history g^(s) = delta^(s) + r^(s)
// re-expressed relative to the new base (the caller then clears alpha', so the
// next z pass rebuilds delta = 0).
__global__ void __launch_bounds__(kThreads) k_rebase(const RhsState* __restrict__ st, long long N,
                                                     float* __restrict__ base, float* __restrict__ delta,
                                                     float* __restrict__ G, const int* __restrict__ head,
                                                     const int* __restrict__ filled,
                                                     const double* __restrict__ alpha_prev, int m1) {
  const int r = blockIdx.y;
  if (!st[r].active) return;
  const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (i >= 6 * N) return;
  const long long off = (long long)r * 6 * N + i;
  if (!G) {
    base[off] += delta[off];
    delta[off] = 0.f;
    return;
  }
  const int f = filled[r], h = head[r];
  float* Gb = G + (long long)r * m1 * 6 * N;
  const long long ss = 6 * N;
  float gv[kMaxDepth1];
  aa_ring_load(Gb, ss, i, h, f, m1, gv);
  const float dl = aa_ring_sum(gv, alpha_prev + (long long)r * m1, h, f, m1, false);
  base[off] += dl;
#pragma unroll
  for (int a = 0; a < kMaxDepth1; ++a)
    if (a < f) Gb[(long long)aa_slot(h, a, m1) * ss + i] = gv[a] - dl;
}

// delta of every selected load case written from the Anderson ring (end of
// a solve, or before the stress of the reference's last-update quirk): with
// the weights of the last z pass for the load cases still active (their last
// step was a z pass), else with the previous weights (their last step was a
// residual evaluation after the ring advanced).
__global__ void __launch_bounds__(kThreads) k_aa_delta(const RhsState* __restrict__ st,
                                                       const int* __restrict__ sel, long long N,
                                                       const float* __restrict__ G, const int* __restrict__ head,
                                                       const int* __restrict__ filled,
                                                       const double* __restrict__ alpha,
                                                       const double* __restrict__ alpha_prev, int m1,
                                                       int last_if_active, float* __restrict__ delta) {
  const int r = blockIdx.y;
  if (sel && !sel[r]) return;
  const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (i >= 6 * N) return;
  const bool cur = last_if_active && st[r].active;
  const double* c = (cur ? alpha : alpha_prev) + (long long)r * m1;
  const int h = head[r], f = filled[r];
  float gv[kMaxDepth1];
  aa_ring_load(G + (long long)r * m1 * 6 * N, 6 * N, i, h, f, m1, gv);
  delta[(long long)r * 6 * N + i] = aa_ring_sum(gv, c, h, f, m1, cur);
}

// alpha' = 0 for the active load cases (after a rebase: delta = 0)
__global__ void k_aa_forget(const RhsState* __restrict__ st, double* __restrict__ alpha_prev, int nrhs, int m1) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrhs || !st[r].active) return;
  for (int s = 0; s < m1; ++s) alpha_prev[(long long)r * m1 + s] = 0.0;
}

// For the load cases with sel[r] != 0: base = src (or eps_bar if src is
// null), delta = 0 (if given).
__global__ void __launch_bounds__(kThreads) k_init_fields(const RhsDesc* __restrict__ rhs, const int* __restrict__ sel,
                                                          long long N, float* __restrict__ base,
                                                          float* __restrict__ delta, const float* __restrict__ src) {
  const int r = blockIdx.y;
  if (sel && !sel[r]) return;
  const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (i >= 6 * N) return;
  const long long off = (long long)r * 6 * N + i;
  const int c = (int)(i / N);
  base[off] = src ? src[off] : rhs[r].eps_bar[c];
  if (delta) delta[off] = 0.f;
}

// ---------------------------------------------------------------------------
// eps = base + delta (optionally written back into delta's buffer),
// sigma = C eps (+tau), optional store, FP64 per-block component sums.
__global__ void __launch_bounds__(kThreads) k_stress(GridDev g, const RhsDesc* __restrict__ rhs,
                                                     const int* __restrict__ sel, const float* __restrict__ lam,
                                                     const float* __restrict__ mu, const float* __restrict__ base,
                                                     float* __restrict__ eps, int write_eps,
                                                     const float* __restrict__ tau, float* __restrict__ sigma,
                                                     double* __restrict__ part6, int chunk) {
  __shared__ double red[kThreads / 64];
  const int r = blockIdx.y;
  if (sel && !sel[r]) return;
  const RhsDesc d = rhs[r];
  const long long N = g.N;
  const float* B_ = base ? base + (long long)r * 6 * N : nullptr;
  float* E_ = eps + (long long)r * 6 * N;
  const float* T_ = tau ? tau + (long long)r * 6 * N : nullptr;
  float* S_ = sigma ? sigma + (long long)r * 6 * N : nullptr;
  const float* LA = lam + (long long)d.mat * N;
  const float* MU = mu + (long long)d.mat * N;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  const long long v0 = (long long)blockIdx.x * chunk;
  const long long v1 = min(N, v0 + chunk);
  for (long long v = v0 + threadIdx.x; v < v1; v += kThreads) {
    float e[6], s[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) e[c] = E_[c * N + v] + (B_ ? B_[c * N + v] : 0.f);
    if (write_eps) {
#pragma unroll
      for (int c = 0; c < 6; ++c) E_[c * N + v] = e[c];
    }
    hooke(e, LA[v], MU[v], s);
    if (g.dmg) {
      double ed[6], sp[6];
#pragma unroll
      for (int c = 0; c < 6; ++c) ed[c] = e[c];
      sigma_tension(ed, (double)LA[v], (double)MU[v], sp);
      const double gm1 = degradation(g.dmg[(long long)d.mat * N + v], g.kdmg) - 1.0;
#pragma unroll
      for (int c = 0; c < 6; ++c) s[c] += (float)(gm1 * sp[c]);
    }
    if (T_) {
#pragma unroll
      for (int c = 0; c < 6; ++c) s[c] += T_[c * N + v];
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      if (S_) S_[c * N + v] = s[c];
      acc[c] += s[c];
    }
  }
  for (int c = 0; c < 6; ++c) {
    const double bs = block_sum(acc[c], red);
    if (threadIdx.x == 0 && part6) part6[((long long)r * gridDim.x + blockIdx.x) * 6 + c] = bs;
  }
}

// Sum the per-block partials of k_stress (fixed order) into out[r][6].
__global__ void __launch_bounds__(kThreads) k_sum6(const double* __restrict__ part6, int nparts,
                                                   double* __restrict__ out, double scale) {
  __shared__ double red[kThreads / 64];
  const int r = blockIdx.x;
  for (int c = 0; c < 6; ++c) {
    double a = 0.0;
    for (int i = threadIdx.x; i < nparts; i += kThreads) a += part6[((long long)r * nparts + i) * 6 + c];
    const double s = block_sum(a, red);
    if (threadIdx.x == 0) out[r * 6 + c] = s * scale;
  }
}

// min/max of the Lame fields of one material (reference medium,
// lippmann_schwinger.cu:172-173): out[mat] = {min lam, max lam, min mu, max mu}
__global__ void __launch_bounds__(kThreads) k_minmax(const float* __restrict__ lam, const float* __restrict__ mu,
                                                     long long N, float* __restrict__ part, int chunk) {
  __shared__ float red[4][kThreads];
  const int mat = blockIdx.y;
  const float* LA = lam + (long long)mat * N;
  const float* MU = mu + (long long)mat * N;
  float a = 3.4e38f, b = -3.4e38f, c = 3.4e38f, dd = -3.4e38f;
  const long long v0 = (long long)blockIdx.x * chunk;
  const long long v1 = min(N, v0 + chunk);
  for (long long v = v0 + threadIdx.x; v < v1; v += kThreads) {
    const float l = LA[v], m = MU[v];
    a = fminf(a, l);
    b = fmaxf(b, l);
    c = fminf(c, m);
    dd = fmaxf(dd, m);
  }
  red[0][threadIdx.x] = a;
  red[1][threadIdx.x] = b;
  red[2][threadIdx.x] = c;
  red[3][threadIdx.x] = dd;
  __syncthreads();
  for (int s = kThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      red[0][threadIdx.x] = fminf(red[0][threadIdx.x], red[0][threadIdx.x + s]);
      red[1][threadIdx.x] = fmaxf(red[1][threadIdx.x], red[1][threadIdx.x + s]);
      red[2][threadIdx.x] = fminf(red[2][threadIdx.x], red[2][threadIdx.x + s]);
      red[3][threadIdx.x] = fmaxf(red[3][threadIdx.x], red[3][threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    float* o = part + ((long long)mat * gridDim.x + blockIdx.x) * 4;
    o[0] = red[0][0];
    o[1] = red[1][0];
    o[2] = red[2][0];
    o[3] = red[3][0];
  }
}

// Gram row of the newest residual of every active load case:
// dots[r][blk][s] = sum over this block's chunk of Re(y_new . conj(y^(s))),
// s = age (0 = newest), FP64 products and sums in a fixed order.  Grid
// (nblk, R), `chunk` complex entries per block.
__global__ void __launch_bounds__(kThreads) k_aa_gram(const RhsState* __restrict__ st, AAParams aa,
                                                      double* __restrict__ dots, long long chunk, int stride) {
  __shared__ double red[kThreads / 64 * kMaxDepth1];
  const int r = blockIdx.y;
  if (!st[r].active) return;
  const int hd = aa.head[r], f = aa.filled[r];
  const float4* base = reinterpret_cast<const float4*>(aa.U + (long long)r * aa.m1 * aa.u_per_rhs);
  const long long slot4 = aa.u_per_rhs / 2;
  const long long i0 = (long long)blockIdx.x * chunk / 2, i1 = min(slot4, i0 + chunk / 2);
  const float4* yn = base + hd * slot4;
  double acc[kMaxDepth1];
#pragma unroll
  for (int q = 0; q < kMaxDepth1; ++q) acc[q] = 0.0;
  for (long long i = i0 + threadIdx.x; i < i1; i += kThreads) {
    const float4 a = ld4nt(&yn[i]);
#pragma unroll
    for (int q = 0; q < kMaxDepth1; ++q) {
      if (q < f) {
        int sl = hd - q;
        if (sl < 0) sl += aa.m1;
        const float4 b = q == 0 ? a : ld4nt(&base[sl * slot4 + i]);
        acc[q] += (double)a.x * b.x + (double)a.y * b.y + (double)a.z * b.z + (double)a.w * b.w;
      }
    }
  }
  block_sum_n(acc, aa.m1, red);
  if (threadIdx.x == 0)
    for (int q = 0; q < aa.m1; ++q) dots[((long long)r * stride + blockIdx.x) * aa.m1 + q] = (q < f) ? acc[q] : 0.0;
}

// Correction of the Gram row for the self-conjugate planes (see AAPlanes):
// dots[r][boff + blk][s] = -scale sum w_c Im R_new Im R^(s), scale = 1/(nx ny)
// (the y ring counts (1/(nx ny)) sum_{x,y} |R|^2 per plane by Parseval).
__global__ void __launch_bounds__(kThreads) k_aa_gram_planes(const RhsState* __restrict__ st, AAPlanes pl,
                                                             const int* __restrict__ filled, double* __restrict__ dots,
                                                             long long chunk, int stride, int boff, double scale) {
  __shared__ double red[kThreads / 64 * kMaxDepth1];
  const int r = blockIdx.y;
  if (!st[r].active) return;
  const int hd = pl.head[r], f = filled[r];
  const float* base = pl.im + (long long)r * pl.m1 * pl.per_slot;
  const long long i0 = (long long)blockIdx.x * chunk, i1 = min(pl.per_slot, i0 + chunk);
  const float* yn = base + hd * pl.per_slot;
  double acc[kMaxDepth1];
#pragma unroll
  for (int q = 0; q < kMaxDepth1; ++q) acc[q] = 0.0;
  for (long long i = i0 + threadIdx.x; i < i1; i += kThreads) {
    const int c = (int)i % 6;  // per_slot < 2^31
    const double a = (c < 3 ? 1.0 : 2.0) * yn[i];
#pragma unroll
    for (int q = 0; q < kMaxDepth1; ++q) {
      if (q < f) {
        int sl = hd - q;
        if (sl < 0) sl += pl.m1;
        acc[q] += a * base[sl * pl.per_slot + i];
      }
    }
  }
  block_sum_n(acc, pl.m1, red);
  if (threadIdx.x == 0)
    for (int q = 0; q < pl.m1; ++q)
      dots[((long long)r * stride + boff + blockIdx.x) * pl.m1 + q] = (q < f) ? -scale * acc[q] : 0.0;
}

// Anderson weights per load case (lippmann_schwinger.py:217-228): reduce the
// new Gram row <r, r^(s)>_w (FP64, fixed order) into row and column 0 of the
// Gram matrix (k_aa_advance has shifted the older ages), solve A v = u on the
// filled block by Gaussian elimination with partial pivoting in FP64,
// alpha = v / (u . v).  A singular or non-finite system falls back to the
// plain step alpha = e_0.  One block per load case.  Running it again after
// the newest row was re-evaluated (a rebase) replaces that row.
__global__ void __launch_bounds__(kThreads) k_aa_solve(const RhsState* __restrict__ st,
                                                       const double* __restrict__ dots, int nblk, int m1,
                                                       double* __restrict__ gram, double* __restrict__ alpha,
                                                       const int* __restrict__ head,
                                                       const int* __restrict__ filled) {
  __shared__ double red[kThreads / 64 * kMaxDepth1];
  const int r = blockIdx.x;
  if (!st[r].active) return;
  const int f = filled[r];
  double row[kMaxDepth1];
#pragma unroll
  for (int s = 0; s < kMaxDepth1; ++s) row[s] = 0.0;
  for (int i = threadIdx.x; i < nblk; i += kThreads) {
    const double* di = dots + ((long long)r * nblk + i) * m1;
#pragma unroll
    for (int s = 0; s < kMaxDepth1; ++s)
      if (s < m1) row[s] += di[s];
  }
  block_sum_n(row, m1, red);
  // the elimination works in LDS (a thread-local array indexed at run time
  // would live in scratch memory); the block updates the stored Gram matrix
  // (new row and column 0, identity on unfilled ages) and loads the filled
  // block into LDS in parallel, then one thread eliminates
  __shared__ double M[kMaxDepth1][kMaxDepth1 + 1];
  __shared__ double v[kMaxDepth1];
  __shared__ double row0[kMaxDepth1];
  if (threadIdx.x == 0) {
#pragma unroll
    for (int s = 0; s < kMaxDepth1; ++s) row0[s] = s < f ? row[s] : 0.0;
  }
  __syncthreads();
  double* A = gram + (long long)r * m1 * m1;
  for (int e = threadIdx.x; e < m1 * m1; e += kThreads) {
    const int i = e / m1, j = e - i * m1;
    double a = i == 0 ? row0[j] : (j == 0 ? row0[i] : A[e]);
    if (i == j && i >= f) a = 1.0;  // unfilled ages keep the identity
    A[e] = a;
    if (i < f && j < f) M[i][j] = a;
  }
  if (threadIdx.x < f) M[threadIdx.x][f] = 1.0;
  __syncthreads();
  if (threadIdx.x != 0) return;
  bool ok = true;
  for (int c = 0; c < f && ok; ++c) {
    int piv = c;
    for (int i = c + 1; i < f; ++i)
      if (fabs(M[i][c]) > fabs(M[piv][c])) piv = i;
    if (!(fabs(M[piv][c]) > 1e-300)) {
      ok = false;
      break;
    }
    if (piv != c)
      for (int j = c; j <= f; ++j) {
        const double t = M[c][j];
        M[c][j] = M[piv][j];
        M[piv][j] = t;
      }
    for (int i = c + 1; i < f; ++i) {
      const double q = M[i][c] / M[c][c];
      for (int j = c; j <= f; ++j) M[i][j] -= q * M[c][j];
    }
  }
  double sv = 0.0;
  if (ok) {
    for (int i = f - 1; i >= 0; --i) {
      double t = M[i][f];
      for (int j = i + 1; j < f; ++j) t -= M[i][j] * v[j];
      v[i] = t / M[i][i];
      sv += v[i];
    }
    ok = (sv == sv) && fabs(sv) > 1e-300 && fabs(sv) < 1e300;
  }
  const int h = head[r];
  double* al = alpha + (long long)r * m1;
  for (int s = 0; s < m1; ++s) al[s] = 0.0;
  for (int s = 0; s < f; ++s) {
    int sl = h - s;
    if (sl < 0) sl += m1;
    al[sl] = ok ? v[s] / sv : (s == 0 ? 1.0 : 0.0);
  }
}

// Advance the Anderson ring of every active load case before the x pass that
// records a new residual: previous weights, newest slot, fill count, and the
// Gram matrix shifted by one age (row and column 0 are then set by
// k_aa_solve).
__global__ void k_aa_advance(const RhsState* __restrict__ st, int* __restrict__ head, int* __restrict__ filled,
                             double* __restrict__ gram, const double* __restrict__ alpha,
                             double* __restrict__ alpha_prev, int nrhs, int m1) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrhs || !st[r].active) return;
  // the weights of the last mix rebuild the current iterate's delta
  for (int s = 0; s < m1; ++s) alpha_prev[(long long)r * m1 + s] = alpha[(long long)r * m1 + s];
  head[r] = (head[r] + 1) % m1;
  filled[r] = min(filled[r] + 1, m1);
  double* A = gram + (long long)r * m1 * m1;
  for (int i = m1 - 1; i >= 1; --i)
    for (int j = m1 - 1; j >= 1; --j) A[i * m1 + j] = A[(i - 1) * m1 + (j - 1)];
}

}  // namespace lsk
