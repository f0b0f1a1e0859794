// Kernels of the learned warm start: a periodic 3-D convolutional
// encoder-decoder evaluated and trained with BF16 MFMA (FP32 accumulation) on
// gfx950.  There is no counterpart in the reference; the network predicts the
// initial strain fluctuation from which the Lippmann-Schwinger iteration starts.
//
// Layouts.  Activations and gradients are bf16, channels last:
// [instance][x][y][z][channel] with a channel stride `ld` (a multiple of 8), so
// a decoder input that concatenates an upsampled tensor and a skip tensor is
// one buffer whose slices are written by their producers.  Master weights are
// fp32 [cout][27][cin] (tap = 9 (dx+1) + 3 (dy+1) + (dz+1)); the kernels read
// packed bf16 copies:
//   forward    Wt[co][k],  k = tap * cin + ci,   rows padded to 16, k to 32
//   data grad  Wd[ci][k'], k' = tap' * cs + co,  Wd[ci][tap'*cs+co] = W[co][26-tap'][ci]
// (cs = channel stride of the incoming gradient).  The data gradient is then
// the forward convolution of the output gradient with Wd.
//
// MFMA.  v_mfma_f32_16x16x32_bf16, D[16x16] += A[16x32] B[32x16] with lane l
// holding A[l%16][8(l/16)..+7], B[8(l/16)..+7][l%16] and
// D[4(l/16)..+3][l%16].  The forward pass computes D = W^T X^T: rows are output
// channels (A = 8 consecutive packed weights), columns are voxels (B = 8
// consecutive channels of one neighbour voxel, one 16-byte load), so each lane
// ends with 4 consecutive channels of one voxel (one 8-byte store).  The
// weight gradient contracts over voxels; both operands are copied to planar
// layouts so a lane reads 8 consecutive voxels of one channel.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include <algorithm>

typedef __attribute__((__vector_size__(8 * sizeof(short)))) short bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

namespace wsk {

constexpr int kThreads = 256;
constexpr float kSlope = 0.1f;  // leaky ReLU

enum : int {
  F_ACT = 1,      // leaky ReLU on the output
  F_DACT = 2,     // multiply by the leaky ReLU derivative read from tsrc
  F_PLANAR = 4,   // fp32 output [instance][cvalid][voxel] (+ addc[instance][c])
};

struct Geo {
  int B, X, Y, Z;
  int nper;  // X*Y*Z
  int nvox;  // B*nper
};

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// float -> bf16, round to nearest even (NaN stays NaN)
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float dact(uint16_t y) { return bf2f(y) > 0.f ? 1.f : kSlope; }

__device__ __forceinline__ void decompose(const Geo& g, int v, int& b, int& x, int& y, int& z) {
  b = v / g.nper;
  int r = v - b * g.nper;
  const int yz = g.Y * g.Z;
  x = r / yz;
  r -= x * yz;
  y = r / g.Z;
  z = r - y * g.Z;
}

__device__ __forceinline__ int wrap(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }

// flat index of the periodic neighbour of (b, x, y, z) at tap
__device__ __forceinline__ int neighbour(const Geo& g, int b, int x, int y, int z, int tap) {
  const int dx = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dz = tap % 3 - 1;
  return ((b * g.X + wrap(x + dx, g.X)) * g.Y + wrap(y + dy, g.Y)) * g.Z + wrap(z + dz, g.Z);
}

__device__ __forceinline__ bf16x8 load8(const uint16_t* p) { return *reinterpret_cast<const bf16x8*>(p); }

// ---------------------------------------------------------------------------
// Convolution, one wave per (CT*16 output channels) x (MT*16 voxels) tile.
template <int CT, int MT>
__global__ void __launch_bounds__(kThreads) k_conv(Geo g, const uint16_t* __restrict__ x, int ldx, int cin,
                                                   const uint16_t* __restrict__ wt, int kpad,
                                                   const float* __restrict__ bias, int cout, void* __restrict__ out,
                                                   int ldo, int flags, const uint16_t* __restrict__ tsrc, int ldt,
                                                   const float* __restrict__ addc, int cvalid) {
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
  const long long vbase = (long long)wave * (16 * MT);
  if (vbase >= g.nvox) return;  // whole wave; no barrier in this kernel
  const int col = lane & 15, grp = lane >> 4;
  int vb[MT], vx[MT], vy[MT], vz[MT];
#pragma unroll
  for (int m = 0; m < MT; ++m) {
    int v = (int)vbase + m * 16 + col;
    v = v < g.nvox ? v : g.nvox - 1;  // tail lanes read a valid voxel, store nothing
    decompose(g, v, vb[m], vx[m], vy[m], vz[m]);
  }
  f32x4 acc[CT][MT];
#pragma unroll
  for (int c = 0; c < CT; ++c)
#pragma unroll
    for (int m = 0; m < MT; ++m) acc[c][m] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int K = 27 * cin;
  for (int kk = 0; kk < kpad; kk += 32) {
    const int k0 = kk + grp * 8;
    bf16x8 bf[MT];
    if (k0 < K) {
      const int tap = k0 / cin;
      const int ci = k0 - tap * cin;
#pragma unroll
      for (int m = 0; m < MT; ++m) {
        const int idx = neighbour(g, vb[m], vx[m], vy[m], vz[m], tap);
        bf[m] = load8(x + (size_t)idx * ldx + ci);
      }
    } else {
#pragma unroll
      for (int m = 0; m < MT; ++m) bf[m] = bf16x8{0, 0, 0, 0, 0, 0, 0, 0};
    }
#pragma unroll
    for (int c = 0; c < CT; ++c) {
      const bf16x8 af = load8(wt + (size_t)(c * 16 + col) * kpad + k0);
#pragma unroll
      for (int m = 0; m < MT; ++m) acc[c][m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, bf[m], acc[c][m], 0, 0, 0);
    }
  }
  // epilogue: lane holds channels co0..co0+3 of voxel v
#pragma unroll
  for (int m = 0; m < MT; ++m) {
    const int v = (int)vbase + m * 16 + col;
    if (v >= g.nvox) continue;
#pragma unroll
    for (int c = 0; c < CT; ++c) {
      const int co0 = c * 16 + grp * 4;
      if (co0 >= cout) continue;
      float r[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        r[i] = acc[c][m][i] + (bias ? bias[co0 + i] : 0.f);
        if (flags & F_ACT) r[i] = r[i] > 0.f ? r[i] : kSlope * r[i];
      }
      if (flags & F_DACT) {
        const uint16_t* t = tsrc + (size_t)v * ldt + co0;
#pragma unroll
        for (int i = 0; i < 4; ++i) r[i] *= dact(t[i]);
      }
      if (flags & F_PLANAR) {
        const int b = v / g.nper, p = v - b * g.nper;
        float* o = reinterpret_cast<float*>(out);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int co = co0 + i;
          if (co < cvalid) o[((size_t)b * cvalid + co) * g.nper + p] = r[i] + (addc ? addc[b * cvalid + co] : 0.f);
        }
      } else {
        uint2 w;
        w.x = (uint32_t)f2bf(r[0]) | ((uint32_t)f2bf(r[1]) << 16);
        w.y = (uint32_t)f2bf(r[2]) | ((uint32_t)f2bf(r[3]) << 16);
        *reinterpret_cast<uint2*>(reinterpret_cast<uint16_t*>(out) + (size_t)v * ldo + co0) = w;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Weight gradient.  Both operands are first copied to planar bf16 [c][nvox]
// (k_planar), so a lane reads 8 consecutive voxels of one channel with one
// 16-byte load; the z offset of a tap is applied by shifting the 8 values by
// one element across two loads (z runs are multiples of 8, periodic).
// part[chunk][co][tap][ci] = sum over the chunk's voxels v of
// dz[v][co] * x[neighbour(v, tap)][ci]; one wave per (chunk, tap, 16-channel
// input tile), all output tiles of the wave accumulated in registers.
constexpr int kWgMaxCo = 128;
constexpr int kWgMaxCi = 192;

// channels-last [nvox][ld] (first c channels) -> planar [c][nvox]
__global__ void __launch_bounds__(kThreads) k_planar(const uint16_t* __restrict__ in, int ld, int c, int nvox,
                                                     uint16_t* __restrict__ out) {
  const int ng = c / 8;
  const long long t = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (t >= (long long)(nvox + 63) / 64 * 64 * ng) return;
  // consecutive threads take consecutive voxels of one channel group, so the
  // 2-byte planar stores of a wave are contiguous
  const long long blk = t / (64LL * ng);
  const int r = (int)(t - blk * 64LL * ng);
  const int cg = r / 64;
  const long long v = blk * 64 + (r % 64);
  if (v >= nvox) return;
  const bf16x8 val = load8(in + (size_t)v * ld + cg * 8);
#pragma unroll
  for (int j = 0; j < 8; ++j) out[(size_t)(cg * 8 + j) * nvox + v] = (uint16_t)val[j];
}

__device__ __forceinline__ bf16x8 shift_in_next(bf16x8 a, bf16x8 b) {
  // elements 1..7 of a, then element 0 of b
  bf16x8 o;
#pragma unroll
  for (int j = 0; j < 7; ++j) o[j] = a[j + 1];
  o[7] = b[0];
  return o;
}

__device__ __forceinline__ bf16x8 shift_in_prev(bf16x8 p, bf16x8 a) {
  // element 7 of p, then elements 0..6 of a
  bf16x8 o;
  o[0] = p[7];
#pragma unroll
  for (int j = 1; j < 8; ++j) o[j] = a[j - 1];
  return o;
}

template <int COT, bool ZRUN8>
__global__ void __launch_bounds__(kThreads) k_wgrad(Geo g, const uint16_t* __restrict__ dzp, int cout,
                                                    const uint16_t* __restrict__ xp, int cin, int ci16, int chunk,
                                                    int nwork, float* __restrict__ part) {
  const int lane = threadIdx.x & 63;
  const int w = blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
  if (w >= nwork) return;  // whole wave; no barrier in this kernel
  const int col = lane & 15, grp = lane >> 4;
  const int bt = w % ci16;
  const int tap = (w / ci16) % 27;
  const int ch = w / (ci16 * 27);
  const int dx = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dz = tap % 3 - 1;
  const int v0 = ch * chunk, v1 = min(v0 + chunk, g.nvox);
  const int ci = bt * 16 + col;
  const bool ci_ok = ci < cin;
  const uint16_t* xrow = xp + (size_t)(ci_ok ? ci : 0) * g.nvox;
  f32x4 acc[COT];
#pragma unroll
  for (int a = 0; a < COT; ++a) acc[a] = f32x4{0.f, 0.f, 0.f, 0.f};
  const bf16x8 zero8 = bf16x8{0, 0, 0, 0, 0, 0, 0, 0};
  for (int vb = v0; vb < v1; vb += 32) {
    const int vs = vb + grp * 8;  // this lane's 8 voxels (nvox % 8 == 0, so all or none are valid)
    const bool v_ok = vs < v1;
    bf16x8 bv = zero8;
    if (ci_ok && v_ok) {
      if (ZRUN8) {
        // the 8 voxels are one segment of a z run (Z % 8 == 0)
        int b, x, y, z;
        decompose(g, vs, b, x, y, z);
        const size_t rowb = (size_t)((b * g.X + wrap(x + dx, g.X)) * g.Y + wrap(y + dy, g.Y)) * g.Z;
        const bf16x8 c0 = load8(xrow + rowb + z);
        if (dz == 0) {
          bv = c0;
        } else if (dz > 0) {
          const int zn = z + 8 == g.Z ? 0 : z + 8;
          bv = shift_in_next(c0, load8(xrow + rowb + zn));
        } else {
          const int zp = z == 0 ? g.Z - 8 : z - 8;
          bv = shift_in_prev(load8(xrow + rowb + zp), c0);
        }
      } else {
        // short z runs (coarse levels): one neighbour lookup per voxel
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          int b, x, y, z;
          decompose(g, vs + j, b, x, y, z);
          bv[j] = (short)xrow[neighbour(g, b, x, y, z, tap)];
        }
      }
    }
#pragma unroll
    for (int a = 0; a < COT; ++a) {
      const int co = a * 16 + col;
      const bf16x8 av = (co < cout && v_ok) ? load8(dzp + (size_t)co * g.nvox + vs) : zero8;
      acc[a] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[a], 0, 0, 0);
    }
  }
  // D rows = output channels a*16 + 4 grp + i, columns = input channel ci
  if (ci_ok) {
#pragma unroll
    for (int a = 0; a < COT; ++a)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int co = a * 16 + grp * 4 + i;
        if (co < cout) part[(((size_t)ch * cout + co) * 27 + tap) * cin + ci] = acc[a][i];
      }
  }
}

// out[i] = sum_c part[c][i] (fixed order)
__global__ void __launch_bounds__(kThreads) k_sum_parts(const float* __restrict__ part, int nparts, long long len,
                                                        float* __restrict__ out) {
  const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (i >= len) return;
  float s = 0.f;
  for (int c = 0; c < nparts; ++c) s += part[(size_t)c * len + i];
  out[i] = s;
}

// bias gradient partials: part[block][c] = sum over the block's voxels of dz[v][c]
__global__ void __launch_bounds__(kThreads) k_bias_parts(const uint16_t* __restrict__ dz, int ldz, int cout,
                                                         int nvox, int chunk, float* __restrict__ part) {
  __shared__ float red[kThreads][8];
  const int ng = (cout + 7) / 8;      // channel groups
  const int rows = kThreads / ng;     // voxels processed in parallel
  const int g8 = threadIdx.x % ng, r = threadIdx.x / ng;
  float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  const int v0 = blockIdx.x * chunk, v1 = min(v0 + chunk, nvox);
  if (r < rows)
    for (int v = v0 + r; v < v1; v += rows) {
      const bf16x8 val = load8(dz + (size_t)v * ldz + g8 * 8);
#pragma unroll
      for (int j = 0; j < 8; ++j) s[j] += bf2f((uint16_t)val[j]);
    }
#pragma unroll
  for (int j = 0; j < 8; ++j) red[threadIdx.x][j] = (r < rows) ? s[j] : 0.f;
  __syncthreads();
  if (threadIdx.x < ng * 8) {
    const int gg = threadIdx.x / 8, j = threadIdx.x % 8;
    float t = 0.f;
    for (int rr = 0; rr < rows; ++rr) t += red[rr * ng + gg][j];
    const int c = gg * 8 + j;
    if (c < cout) part[(size_t)blockIdx.x * cout + c] = t;
  }
}

// ---------------------------------------------------------------------------
// 2x average pooling (fine geometry g), thread per (coarse voxel, 8 channels)
__global__ void __launch_bounds__(kThreads) k_pool(Geo g, const uint16_t* __restrict__ in, int ldi, int c,
                                                   uint16_t* __restrict__ out, int ldo) {
  const int ng = c / 8;
  const int X2 = g.X / 2, Y2 = g.Y / 2, Z2 = g.Z / 2;
  const long long n2 = (long long)g.B * X2 * Y2 * Z2;
  const long long t = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (t >= n2 * ng) return;
  const int cg = (int)(t % ng);
  const long long vc = t / ng;
  int r = (int)vc;
  const int z = r % Z2; r /= Z2;
  const int y = r % Y2; r /= Y2;
  const int x = r % X2;
  const int b = r / X2;
  float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  for (int d = 0; d < 8; ++d) {
    const int fi = ((b * g.X + 2 * x + (d >> 2)) * g.Y + 2 * y + ((d >> 1) & 1)) * g.Z + 2 * z + (d & 1);
    const bf16x8 val = load8(in + (size_t)fi * ldi + cg * 8);
#pragma unroll
    for (int j = 0; j < 8; ++j) s[j] += bf2f((uint16_t)val[j]);
  }
  bf16x8 o;
#pragma unroll
  for (int j = 0; j < 8; ++j) o[j] = (short)f2bf(0.125f * s[j]);
  *reinterpret_cast<bf16x8*>(out + (size_t)vc * ldo + cg * 8) = o;
}

// gradient of a skip tensor s (fine geometry g) that feeds both a pooling
// and a decoder concatenation:
//   dz[v] = act'(s[v]) * dcoarse[parent(v)] / 8 + (dcat ? dcat[v] : 0)
// (dcat already carries act'(s), written by the decoder's data-gradient pass)
__global__ void __launch_bounds__(kThreads) k_pool_back(Geo g, const uint16_t* __restrict__ dco, int ldc, int c,
                                                        const uint16_t* __restrict__ s, int lds,
                                                        const uint16_t* __restrict__ dcat, int ldcat,
                                                        uint16_t* __restrict__ dz, int ldz) {
  const int ng = c / 8;
  const long long t = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (t >= (long long)g.nvox * ng) return;
  const int cg = (int)(t % ng);
  const int v = (int)(t / ng);
  int b, x, y, z;
  decompose(g, v, b, x, y, z);
  const int X2 = g.X / 2, Y2 = g.Y / 2, Z2 = g.Z / 2;
  const int pc = ((b * X2 + x / 2) * Y2 + y / 2) * Z2 + z / 2;
  const bf16x8 dv = load8(dco + (size_t)pc * ldc + cg * 8);
  const bf16x8 sv = load8(s + (size_t)v * lds + cg * 8);
  bf16x8 cv = bf16x8{0, 0, 0, 0, 0, 0, 0, 0};
  if (dcat) cv = load8(dcat + (size_t)v * ldcat + cg * 8);
  bf16x8 o;
#pragma unroll
  for (int j = 0; j < 8; ++j)
    o[j] = (short)f2bf(dact((uint16_t)sv[j]) * 0.125f * bf2f((uint16_t)dv[j]) + bf2f((uint16_t)cv[j]));
  *reinterpret_cast<bf16x8*>(dz + (size_t)v * ldz + cg * 8) = o;
}

// nearest 2x upsampling into a (slice of a) fine tensor; g = fine geometry
__global__ void __launch_bounds__(kThreads) k_upsample(Geo g, const uint16_t* __restrict__ in, int ldi, int c,
                                                       uint16_t* __restrict__ out, int ldo) {
  const int ng = c / 8;
  const long long t = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (t >= (long long)g.nvox * ng) return;
  const int cg = (int)(t % ng);
  const int v = (int)(t / ng);
  int b, x, y, z;
  decompose(g, v, b, x, y, z);
  const int pc = ((b * (g.X / 2) + x / 2) * (g.Y / 2) + y / 2) * (g.Z / 2) + z / 2;
  *reinterpret_cast<bf16x8*>(out + (size_t)v * ldo + cg * 8) = load8(in + (size_t)pc * ldi + cg * 8);
}

// gradient of the upsampling: dcoarse[p] = sum of the 8 children; g = fine geometry
__global__ void __launch_bounds__(kThreads) k_upsample_back(Geo g, const uint16_t* __restrict__ dfine, int ldf,
                                                            int c, uint16_t* __restrict__ dco, int ldc) {
  const int ng = c / 8;
  const int X2 = g.X / 2, Y2 = g.Y / 2, Z2 = g.Z / 2;
  const long long n2 = (long long)g.B * X2 * Y2 * Z2;
  const long long t = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (t >= n2 * ng) return;
  const int cg = (int)(t % ng);
  const long long vc = t / ng;
  int r = (int)vc;
  const int z = r % Z2; r /= Z2;
  const int y = r % Y2; r /= Y2;
  const int x = r % X2;
  const int b = r / X2;
  float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  for (int d = 0; d < 8; ++d) {
    const int fi = ((b * g.X + 2 * x + (d >> 2)) * g.Y + 2 * y + ((d >> 1) & 1)) * g.Z + 2 * z + (d & 1);
    const bf16x8 val = load8(dfine + (size_t)fi * ldf + cg * 8);
#pragma unroll
    for (int j = 0; j < 8; ++j) s[j] += bf2f((uint16_t)val[j]);
  }
  bf16x8 o;
#pragma unroll
  for (int j = 0; j < 8; ++j) o[j] = (short)f2bf(s[j]);
  *reinterpret_cast<bf16x8*>(dco + (size_t)vc * ldc + cg * 8) = o;
}

// ---------------------------------------------------------------------------
// network input [instance][voxel][8] bf16: (log lambda - a0)/s0, (log mu - a1)/s1,
// eps_bar[instance][0..5]; instance b uses material b / per_mat
__global__ void __launch_bounds__(kThreads) k_input(const float* __restrict__ lam, const float* __restrict__ mu,
                                                    int per_mat, const float* __restrict__ eps_bar, int B,
                                                    long long nper, float a0, float s0, float a1, float s1,
                                                    uint16_t* __restrict__ out) {
  const long long t = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (t >= (long long)B * nper) return;
  const int b = (int)(t / nper);
  const long long p = t - (long long)b * nper;
  const long long mi = (long long)(b / per_mat) * nper + p;
  bf16x8 o;
  o[0] = (short)f2bf((logf(lam[mi]) - a0) * s0);
  o[1] = (short)f2bf((logf(mu[mi]) - a1) * s1);
#pragma unroll
  for (int j = 0; j < 6; ++j) o[2 + j] = (short)f2bf(eps_bar[b * 6 + j]);
  *reinterpret_cast<bf16x8*>(out + (size_t)t * 8) = o;
}

// training target: bf16 planar fluctuation eps - eps_bar, [B][6][nper]
__global__ void __launch_bounds__(kThreads) k_target(const float* __restrict__ eps, const float* __restrict__ eps_bar,
                                                     int B, long long nper, uint16_t* __restrict__ out) {
  const long long t = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (t >= (long long)B * 6 * nper) return;
  const int bc = (int)(t / nper);
  out[t] = f2bf(eps[t] - eps_bar[bc]);
}

// mean squared error of pred (fp32 planar [B][6][nper]) against target (bf16
// planar); writes the gradient bf16 channels last [voxel][8] (channels 6, 7
// zero) and FP64 block partial sums of the squared error
__global__ void __launch_bounds__(kThreads) k_mse(const float* __restrict__ pred, const uint16_t* __restrict__ tgt,
                                                  int B, long long nper, float gscale, uint16_t* __restrict__ grad,
                                                  double* __restrict__ part) {
  __shared__ double red[kThreads];
  const long long v = (long long)blockIdx.x * kThreads + threadIdx.x;
  double e = 0.0;
  if (v < (long long)B * nper) {
    const int b = (int)(v / nper);
    const long long p = v - (long long)b * nper;
    bf16x8 o = bf16x8{0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      const size_t i = ((size_t)b * 6 + c) * nper + p;
      const float d = pred[i] - bf2f(tgt[i]);
      e += (double)d * d;
      o[c] = (short)f2bf(gscale * d);
    }
    *reinterpret_cast<bf16x8*>(grad + (size_t)v * 8) = o;
  }
  red[threadIdx.x] = e;
  __syncthreads();
  for (int s = kThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// ---------------------------------------------------------------------------
// Training loss on the solver's own terms (ls_ws_residual_loss).  Per sample
// s and unit load case k (K = 6) with the prediction error e = pred - target
// (the fluctuation, six Voigt components):
//   L = sum_{s,k} [ wm <e, C e>_w / <t, C t>_w + wr |D^- . C e|^2 / |D^- . C e_k|^2 ] / (S K),
// C the sample's Hooke law, <., .>_w the Voigt-weighted product (so <e, C e>_w
// is the strain energy of the error: the norm conjugate gradients reduce), D^-
// the backward divergence of the rotated staggered grid (the 2 x 2 x 2 cell of
// voxels (i - a, j - b, k - c)).  The target is an equilibrium field
// (D^- . C (e_k + t) = 0 to the solver's tolerance), so the second term is the
// squared equilibrium residual of the predicted strain relative to the uniform
// start's: the solver's stop measure (equal to its Fourier residual by
// Parseval).  Partials are FP64 per block in a fixed order.  Channel index
// ch = 6 k + c.
constexpr int kRlK = 6;
struct RlGeo {
  int S, X, Y, Z;
  long long nper;
  float sx, sy, sz;  // n_d / (4 L_d)
};

__device__ __forceinline__ void rl_ijk(const RlGeo& g, long long v, int& i, int& j, int& k) {
  k = (int)(v % g.Z);
  const long long q = v / g.Z;
  j = (int)(q % g.Y);
  i = (int)(q / g.Y);
}

// stress of the error, sigma_e = C e, into sc [S][36][nper]; partials of <e, C e>_w and <t, C t>_w per k
__global__ void __launch_bounds__(kThreads) k_rl_stress(RlGeo g, const float* __restrict__ pred,
                                                        const uint16_t* __restrict__ tgt,
                                                        const float* __restrict__ lam, const float* __restrict__ mu,
                                                        float* __restrict__ sc, double* __restrict__ part) {
  __shared__ double red[2 * kRlK][kThreads / 64];
  const int s = blockIdx.y;
  const long long v = (long long)blockIdx.x * kThreads + threadIdx.x;
  double acc[2 * kRlK];
#pragma unroll
  for (int q = 0; q < 2 * kRlK; ++q) acc[q] = 0.0;
  if (v < g.nper) {
    const float l = lam[(long long)s * g.nper + v], m2 = 2.f * mu[(long long)s * g.nper + v];
#pragma unroll
    for (int k = 0; k < kRlK; ++k) {
      float e[6], t[6];
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        const size_t idx = ((size_t)s * 36 + 6 * k + c) * g.nper + v;
        t[c] = bf2f(tgt[idx]);
        e[c] = pred[idx] - t[c];
      }
      const float tr = e[0] + e[1] + e[2], trt = t[0] + t[1] + t[2];
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        const float se = m2 * e[c] + (c < 3 ? l * tr : 0.f), st = m2 * t[c] + (c < 3 ? l * trt : 0.f);
        const double wc = c < 3 ? 1.0 : 2.0;
        acc[2 * k] += wc * (double)e[c] * se;
        acc[2 * k + 1] += wc * (double)t[c] * st;
        sc[((size_t)s * 36 + 6 * k + c) * g.nper + v] = se;
      }
    }
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < 2 * kRlK; ++q) {
    double a = acc[q];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
    if (lane == 0) red[q][w] = a;
  }
  __syncthreads();
  if (threadIdx.x < 2 * kRlK) {
    double a = 0.0;
    for (int i = 0; i < kThreads / 64; ++i) a += red[threadIdx.x][i];
    part[((size_t)s * gridDim.x + blockIdx.x) * (4 * kRlK) + threadIdx.x] = a;
  }
}

// divergence of the error stress, dv = D^- . sigma_e [S][K][3][nper], and of the
// uniform start's stress C e_k; partials of |dv|^2 and |D^- . C e_k|^2 per k
__global__ void __launch_bounds__(kThreads) k_rl_div(RlGeo g, const float* __restrict__ sc,
                                                     const float* __restrict__ lam, const float* __restrict__ mu,
                                                     float* __restrict__ dv, double* __restrict__ part) {
  __shared__ double red[2 * kRlK][kThreads / 64];
  const int s = blockIdx.y;
  const long long v = (long long)blockIdx.x * kThreads + threadIdx.x;
  double acc[2 * kRlK];
#pragma unroll
  for (int q = 0; q < 2 * kRlK; ++q) acc[q] = 0.0;
  if (v < g.nper) {
    int i, j, k;
    rl_ijk(g, v, i, j, k);
    const int im = i == 0 ? g.X - 1 : i - 1, jm = j == 0 ? g.Y - 1 : j - 1, km = k == 0 ? g.Z - 1 : k - 1;
    const long long ii[2] = {(long long)i * g.Y * g.Z, (long long)im * g.Y * g.Z};
    const long long jj[2] = {(long long)j * g.Z, (long long)jm * g.Z};
    const int kk[2] = {k, km};
    // D^-_d of one field f (offset base) over the cell
    auto dd = [&](const float* f, float& dx, float& dy, float& dz) {
      float c[2][2][2];
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
          for (int e = 0; e < 2; ++e) c[a][b][e] = f[ii[a] + jj[b] + kk[e]];
      float tx = 0.f, ty = 0.f, tz = 0.f;
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
          tx += c[0][a][b] - c[1][a][b];
          ty += c[a][0][b] - c[a][1][b];
          tz += c[a][b][0] - c[a][b][1];
        }
      dx = g.sx * tx;
      dy = g.sy * ty;
      dz = g.sz * tz;
    };
    const float* L = lam + (long long)s * g.nper;
    const float* Mu = mu + (long long)s * g.nper;
    float lx, ly, lz, mx, my, mz;
    dd(L, lx, ly, lz);
    dd(Mu, mx, my, mz);
#pragma unroll
    for (int q = 0; q < kRlK; ++q) {
      const float* base = sc + ((size_t)s * 36 + 6 * q) * g.nper;
      float d0[6][3];
#pragma unroll
      for (int c = 0; c < 6; ++c) dd(base + (size_t)c * g.nper, d0[c][0], d0[c][1], d0[c][2]);
      // div_x = D_x s_xx + D_y s_xy + D_z s_xz, div_y = D_x s_xy + D_y s_yy + D_z s_yz,
      // div_z = D_x s_xz + D_y s_yz + D_z s_zz (Voigt xx yy zz xy xz yz)
      const float q0 = d0[0][0] + d0[3][1] + d0[4][2];
      const float q1 = d0[3][0] + d0[1][1] + d0[5][2];
      const float q2 = d0[4][0] + d0[5][1] + d0[2][2];
      float* o = dv + ((size_t)s * kRlK + q) * 3 * g.nper + v;
      o[0] = q0;
      o[g.nper] = q1;
      o[2 * g.nper] = q2;
      acc[2 * q] += (double)q0 * q0 + (double)q1 * q1 + (double)q2 * q2;
      // the uniform start's stress C e_q: 2 mu e_q + lambda tr(e_q) I
      float u0, u1, u2;
      if (q < 3) {
        const float dl[3] = {lx, ly, lz}, dm[3] = {mx, my, mz};
        u0 = dl[0] + (q == 0 ? 2.f * dm[0] : 0.f);
        u1 = dl[1] + (q == 1 ? 2.f * dm[1] : 0.f);
        u2 = dl[2] + (q == 2 ? 2.f * dm[2] : 0.f);
      } else {
        // shear e_q (tensor component 1 in slot q): sigma_q = 2 mu
        const float dm[3] = {mx, my, mz};
        const int a = q == 3 ? 0 : (q == 4 ? 0 : 1), b = q == 3 ? 1 : (q == 4 ? 2 : 2);
        float w3[3] = {0.f, 0.f, 0.f};
        w3[a] = 2.f * dm[b];
        w3[b] = 2.f * dm[a];
        u0 = w3[0];
        u1 = w3[1];
        u2 = w3[2];
      }
      acc[2 * q + 1] += (double)u0 * u0 + (double)u1 * u1 + (double)u2 * u2;
    }
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < 2 * kRlK; ++q) {
    double a = acc[q];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
    if (lane == 0) red[q][w] = a;
  }
  __syncthreads();
  if (threadIdx.x < 2 * kRlK) {
    double a = 0.0;
    for (int i = 0; i < kThreads / 64; ++i) a += red[threadIdx.x][i];
    part[((size_t)s * gridDim.x + blockIdx.x) * (4 * kRlK) + 2 * kRlK + threadIdx.x] = a;
  }
}

// per (s, k): the two scale factors of the gradient, and the loss (one block)
__global__ void __launch_bounds__(kThreads) k_rl_reduce(int S, int nblk, float wm, float wr,
                                                        const double* __restrict__ part,
                                                        float* __restrict__ scale, double* __restrict__ loss) {
  __shared__ double sums[64][4 * kRlK];
  __shared__ double tot;
  const int nq = 4 * kRlK;
  if (threadIdx.x == 0) tot = 0.0;
  __syncthreads();
  for (int s0 = 0; s0 < S; s0 += 64) {
    const int ns = min(64, S - s0);
    for (int t = threadIdx.x; t < ns * nq; t += kThreads) {
      const int sl = t / nq, q = t - sl * nq;
      double a = 0.0;
      for (int b = 0; b < nblk; ++b) a += part[((size_t)(s0 + sl) * nblk + b) * nq + q];
      sums[sl][q] = a;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      const double inv = 1.0 / ((double)S * kRlK);
      for (int sl = 0; sl < ns; ++sl)
        for (int k = 0; k < kRlK; ++k) {
          const double E = sums[sl][2 * k], T = fmax(sums[sl][2 * k + 1], 1e-30);
          const double Q = sums[sl][2 * kRlK + 2 * k], D = fmax(sums[sl][2 * kRlK + 2 * k + 1], 1e-30);
          tot += inv * (wm * E / T + wr * Q / D);
          scale[((s0 + sl) * kRlK + k) * 2] = (float)(2.0 * wm * inv / T);
          scale[((s0 + sl) * kRlK + k) * 2 + 1] = (float)(2.0 * wr * inv / D);
        }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) loss[0] = tot;
}

// dL/dpred = cm w C e + cr C (D^-)^T dv (w the Voigt weights), with (D^-_x)^T f (x) = sx sum_{b,c} f(i, j + b, k + c) - f(i + 1, j + b, k + c)
// (and likewise along y and z); into grad bf16 [S][nper][ldg], channels >= 36 zero
__global__ void __launch_bounds__(kThreads) k_rl_grad(RlGeo g, const float* __restrict__ sc,
                                                      const float* __restrict__ lam, const float* __restrict__ mu,
                                                      const float* __restrict__ dv, const float* __restrict__ scale,
                                                      uint16_t* __restrict__ grad, int ldg) {
  const int s = blockIdx.y;
  const long long v = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (v >= g.nper) return;
  int i, j, k;
  rl_ijk(g, v, i, j, k);
  const int ip = i + 1 == g.X ? 0 : i + 1, jp = j + 1 == g.Y ? 0 : j + 1, kp = k + 1 == g.Z ? 0 : k + 1;
  const long long ii[2] = {(long long)i * g.Y * g.Z, (long long)ip * g.Y * g.Z};
  const long long jj[2] = {(long long)j * g.Z, (long long)jp * g.Z};
  const int kk[2] = {k, kp};
  auto dt = [&](const float* f, float& dx, float& dy, float& dz) {
    float c[2][2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int e = 0; e < 2; ++e) c[a][b][e] = f[ii[a] + jj[b] + kk[e]];
    float tx = 0.f, ty = 0.f, tz = 0.f;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        tx += c[0][a][b] - c[1][a][b];
        ty += c[a][0][b] - c[a][1][b];
        tz += c[a][b][0] - c[a][b][1];
      }
    dx = g.sx * tx;
    dy = g.sy * ty;
    dz = g.sz * tz;
  };
  const float l = lam[(long long)s * g.nper + v], m2 = 2.f * mu[(long long)s * g.nper + v];
  uint16_t* o = grad + ((size_t)s * g.nper + v) * ldg;
  for (int q = 0; q < kRlK; ++q) {
    const float* f = dv + ((size_t)s * kRlK + q) * 3 * g.nper;
    float a[3][3];  // a[i][d] = (D^-_d)^T dv_i
#pragma unroll
    for (int c = 0; c < 3; ++c) dt(f + (size_t)c * g.nper, a[c][0], a[c][1], a[c][2]);
    // d/dsigma (Voigt): xx: a_x,x; yy: a_y,y; zz: a_z,z; xy: a_x,y + a_y,x; xz: a_x,z + a_z,x; yz: a_y,z + a_z,y
    float gs[6] = {a[0][0], a[1][1], a[2][2], a[0][1] + a[1][0], a[0][2] + a[2][0], a[1][2] + a[2][1]};
    const float trg = gs[0] + gs[1] + gs[2];
    const float cm = scale[(s * kRlK + q) * 2], cr = scale[(s * kRlK + q) * 2 + 1];
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      const float se = sc[((size_t)s * 36 + 6 * q + c) * g.nper + v];  // (C e)_c
      const float ge = m2 * gs[c] + (c < 3 ? l * trg : 0.f);          // C (symmetric) applied to the stress gradient
      o[6 * q + c] = f2bf(cm * (c < 3 ? 1.f : 2.f) * se + cr * ge);
    }
  }
  for (int c = 36; c < ldg; ++c) o[c] = 0;
}

// ---------------------------------------------------------------------------
// Compatible strains from predicted displacements (ls_ws_symgrad): per load
// case k, eps_k = sym(D^+ u_k) with D^+_d = -(D^-_d)^T the forward derivative
// of the rotated staggered grid,
//   D^+_x f(i, j, k) = sx sum_{b,c in {0,1}} f(i + 1, j + b, k + c) - f(i, j + b, k + c)
// (and likewise along y, z), Voigt xx yy zz xy xz yz with tensor shear.  In
// Fourier D^+_d is i xi~_d times a phase common to the three axes, so
// sym(D^+ u) lies in the range of the solver's Gamma0: the start needs no
// projection.  u fp32 [S][18][nper] (3 k + c), eps fp32 [S][36][nper]
// (6 k + c), plus e_k if add_unit.
__global__ void __launch_bounds__(kThreads) k_symgrad(RlGeo g, const float* __restrict__ u, int add_unit,
                                                      float* __restrict__ eps) {
  const int s = blockIdx.y;
  const long long v = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (v >= g.nper) return;
  int i, j, k;
  rl_ijk(g, v, i, j, k);
  const int ip = i + 1 == g.X ? 0 : i + 1, jp = j + 1 == g.Y ? 0 : j + 1, kp = k + 1 == g.Z ? 0 : k + 1;
  const long long ii[2] = {(long long)i * g.Y * g.Z, (long long)ip * g.Y * g.Z};
  const long long jj[2] = {(long long)j * g.Z, (long long)jp * g.Z};
  const int kk[2] = {k, kp};
  for (int q = 0; q < kRlK; ++q) {
    float d[3][3];  // d[c][dir] = D^+_dir u_c
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float* f = u + ((size_t)s * 18 + 3 * q + c) * g.nper;
      float cell[2][2][2];
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
          for (int e = 0; e < 2; ++e) cell[a][b][e] = f[ii[a] + jj[b] + kk[e]];
      float tx = 0.f, ty = 0.f, tz = 0.f;
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
          tx += cell[1][a][b] - cell[0][a][b];
          ty += cell[a][1][b] - cell[a][0][b];
          tz += cell[a][b][1] - cell[a][b][0];
        }
      d[c][0] = g.sx * tx;
      d[c][1] = g.sy * ty;
      d[c][2] = g.sz * tz;
    }
    float e[6] = {d[0][0], d[1][1], d[2][2], 0.5f * (d[0][1] + d[1][0]), 0.5f * (d[0][2] + d[2][0]),
                  0.5f * (d[1][2] + d[2][1])};
    if (add_unit) e[q] += 1.f;
#pragma unroll
    for (int c = 0; c < 6; ++c) eps[((size_t)s * 36 + 6 * q + c) * g.nper + v] = e[c];
  }
}

// Adjoint: dL/du from dL/deps (bf16 [S][nper][ldg], channel 6 k + c), with
// (D^+_d)^T = -D^-_d: dL/du_x = -(D^-_x g_xx + D^-_y g_xy / 2 + D^-_z g_xz / 2), etc.
// Into bf16 [S][nper][ldu] (channel 3 k + c; channels >= 18 zero).
__global__ void __launch_bounds__(kThreads) k_symgrad_back(RlGeo g, const uint16_t* __restrict__ ge, int ldg,
                                                           uint16_t* __restrict__ gu, int ldu) {
  const int s = blockIdx.y;
  const long long v = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (v >= g.nper) return;
  int i, j, k;
  rl_ijk(g, v, i, j, k);
  const int im = i == 0 ? g.X - 1 : i - 1, jm = j == 0 ? g.Y - 1 : j - 1, km = k == 0 ? g.Z - 1 : k - 1;
  const long long ii[2] = {(long long)i * g.Y * g.Z, (long long)im * g.Y * g.Z};
  const long long jj[2] = {(long long)j * g.Z, (long long)jm * g.Z};
  const int kk[2] = {k, km};
  const uint16_t* G = ge + (size_t)s * g.nper * ldg;
  uint16_t* o = gu + ((size_t)s * g.nper + v) * ldu;
  for (int q = 0; q < kRlK; ++q) {
    float dm[6][3];  // D^-_dir of the six components of dL/deps_k
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      float cell[2][2][2];
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
          for (int e = 0; e < 2; ++e) cell[a][b][e] = bf2f(G[(size_t)(ii[a] + jj[b] + kk[e]) * ldg + 6 * q + c]);
      float tx = 0.f, ty = 0.f, tz = 0.f;
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
          tx += cell[0][a][b] - cell[1][a][b];
          ty += cell[a][0][b] - cell[a][1][b];
          tz += cell[a][b][0] - cell[a][b][1];
        }
      dm[c][0] = g.sx * tx;
      dm[c][1] = g.sy * ty;
      dm[c][2] = g.sz * tz;
    }
    const float ux = -(dm[0][0] + 0.5f * dm[3][1] + 0.5f * dm[4][2]);
    const float uy = -(0.5f * dm[3][0] + dm[1][1] + 0.5f * dm[5][2]);
    const float uz = -(0.5f * dm[4][0] + 0.5f * dm[5][1] + dm[2][2]);
    o[3 * q] = f2bf(ux);
    o[3 * q + 1] = f2bf(uy);
    o[3 * q + 2] = f2bf(uz);
  }
  for (int c = 18; c < ldu; ++c) o[c] = 0;
}

__global__ void __launch_bounds__(kThreads) k_sum_f64(const double* __restrict__ part, int n, double* __restrict__ out) {
  __shared__ double red[kThreads];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += kThreads) s += part[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = kThreads / 2; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

// Adam (Kingma & Ba 2015) on fp32 master parameters
__global__ void __launch_bounds__(kThreads) k_adam(float* __restrict__ w, const float* __restrict__ g,
                                                   float* __restrict__ m, float* __restrict__ v, long long n, float lr,
                                                   float b1, float b2, float eps, float c1, float c2) {
  const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

// pack fp32 W[cout][27][cin] into the bf16 forward (Wt) and data-gradient (Wd) layouts
__global__ void __launch_bounds__(kThreads) k_pack(const float* __restrict__ w, int cout, int cin, int cs,
                                                   uint16_t* __restrict__ wt, int kpad, int rows_t,
                                                   uint16_t* __restrict__ wd, int kpad_d, int rows_d) {
  const long long t = (long long)blockIdx.x * kThreads + threadIdx.x;
  const long long nt = (long long)rows_t * kpad;
  const long long nd = wd ? (long long)rows_d * kpad_d : 0;
  if (t < nt) {
    const int co = (int)(t / kpad), k = (int)(t % kpad);
    float val = 0.f;
    if (co < cout && k < 27 * cin) {
      const int tap = k / cin, ci = k % cin;
      val = w[((size_t)co * 27 + tap) * cin + ci];
    }
    wt[t] = f2bf(val);
  } else if (t < nt + nd) {
    const long long u = t - nt;
    const int ci = (int)(u / kpad_d), k = (int)(u % kpad_d);
    float val = 0.f;
    if (ci < cin && k < 27 * cs) {
      const int tp = k / cs, co = k % cs;
      if (co < cout) val = w[((size_t)co * 27 + (26 - tp)) * cin + ci];
    }
    wd[u] = f2bf(val);
  }
}

bool rl_geo(const int* dims, const float* extents, RlGeo& g) {
  if (!dims || !extents || dims[0] < 1 || dims[1] < 1 || dims[2] < 1 || dims[3] < 1 ||
      !(extents[0] > 0.f && extents[1] > 0.f && extents[2] > 0.f))
    return false;
  g.S = dims[0];
  g.X = dims[1];
  g.Y = dims[2];
  g.Z = dims[3];
  g.nper = (long long)g.X * g.Y * g.Z;
  g.sx = (float)(0.25 * g.X / (double)extents[0]);
  g.sy = (float)(0.25 * g.Y / (double)extents[1]);
  g.sz = (float)(0.25 * g.Z / (double)extents[2]);
  return true;
}

}  // namespace wsk

// ============================================================================
// C ABI
using namespace wsk;

namespace {

int launch_ok(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    fprintf(stderr, "ls_ws: %s: %s\n", what, hipGetErrorString(e));
    return -1;
  }
  return 0;
}

bool bad_geo(const int* dims) {
  return !dims || dims[0] < 1 || dims[1] < 1 || dims[2] < 1 || dims[3] < 1 ||
         (long long)dims[0] * dims[1] * dims[2] * dims[3] >= (1LL << 31);
}

Geo make_geo(const int* dims) {
  Geo g;
  g.B = dims[0];
  g.X = dims[1];
  g.Y = dims[2];
  g.Z = dims[3];
  g.nper = g.X * g.Y * g.Z;
  g.nvox = g.B * g.nper;
  return g;
}

unsigned nblocks(long long n) { return (unsigned)((n + kThreads - 1) / kThreads); }

bool misaligned8(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) != 0; }

}  // namespace

extern "C" {

int ls_ws_conv3d(const int* dims, const uint16_t* x, int ldx, int cin, const uint16_t* wt, int kpad,
                 const float* bias, int cout, void* out, int ldo, int flags, const uint16_t* tsrc, int ldt,
                 const float* addc, int cvalid, void* stream) {
  if (bad_geo(dims) || !x || !wt || !out || cin < 8 || cin % 8 || ldx % 8 || ldx < cin || cout < 1 || cout > 192 ||
      cout % 4 || kpad % 32 || kpad < 27 * cin || misaligned8(x) || misaligned8(wt)) {
    fprintf(stderr, "ls_ws_conv3d: invalid arguments (cin %d ldx %d cout %d kpad %d)\n", cin, ldx, cout, kpad);
    return -1;
  }
  if (!(flags & F_PLANAR) && (ldo % 4 || ldo < cout || (reinterpret_cast<uintptr_t>(out) & 7u))) {
    fprintf(stderr, "ls_ws_conv3d: invalid output stride %d\n", ldo);
    return -1;
  }
  if ((flags & F_DACT) && (!tsrc || ldt < cout)) {
    fprintf(stderr, "ls_ws_conv3d: F_DACT needs the forward activation\n");
    return -1;
  }
  if ((flags & F_PLANAR) && (cvalid < 1 || cvalid > cout)) {
    fprintf(stderr, "ls_ws_conv3d: invalid cvalid %d\n", cvalid);
    return -1;
  }
  const Geo g = make_geo(dims);
  hipStream_t st = (hipStream_t)stream;
  // one instantiation per tile height: the packed weights have exactly 16 * ceil(cout / 16) rows
  const int ct = (cout + 15) / 16;
#define LS_WS_CONV(CT, MT)                                                                                     \
  case CT: {                                                                                                   \
    const long long waves = ((long long)g.nvox + 16 * MT - 1) / (16 * MT);                                     \
    k_conv<CT, MT><<<(unsigned)((waves + 3) / 4), kThreads, 0, st>>>(g, x, ldx, cin, wt, kpad, bias, cout, out,  \
                                                                     ldo, flags, tsrc, ldt, addc, cvalid);     \
    break;                                                                                                     \
  }
  switch (ct) {
    LS_WS_CONV(1, 8)
    LS_WS_CONV(2, 8)
    LS_WS_CONV(3, 4)
    LS_WS_CONV(4, 4)
    LS_WS_CONV(6, 2)
    LS_WS_CONV(8, 2)
    LS_WS_CONV(12, 1)
    default:
      fprintf(stderr, "ls_ws_conv3d: unsupported output channel count %d\n", cout);
      return -1;
  }
#undef LS_WS_CONV
  return launch_ok("conv3d");
}

int ls_ws_conv3d_wgrad(const int* dims, const uint16_t* dz, int ldz, int cout, const uint16_t* x, int ldx, int cin,
                       uint16_t* planar, long long planar_cap, float* part, long long part_cap, float* dw, float* db,
                       void* stream) {
  if (bad_geo(dims) || !dz || !x || !planar || !part || !dw || cout < 1 || cout > kWgMaxCo || cout % 8 || cin < 8 ||
      cin > kWgMaxCi || cin % 8 || ldx % 8 || ldz % 8 || ldz < cout || ldx < cin || misaligned8(dz) ||
      misaligned8(x) || misaligned8(planar)) {
    fprintf(stderr, "ls_ws_conv3d_wgrad: invalid arguments (cout %d cin %d nz %d)\n", cout, cin, dims ? dims[3] : 0);
    return -1;
  }
  const Geo g = make_geo(dims);
  if (g.nvox % 8) {
    fprintf(stderr, "ls_ws_conv3d_wgrad: the voxel count must be a multiple of 8\n");
    return -1;
  }
  if ((long long)(cout + cin) * g.nvox > planar_cap) {
    fprintf(stderr, "ls_ws_conv3d_wgrad: planar buffer too small\n");
    return -1;
  }
  hipStream_t st = (hipStream_t)stream;
  uint16_t* dzp = planar;
  uint16_t* xp = planar + (size_t)cout * g.nvox;
  const long long nv64 = (g.nvox + 63) / 64 * 64;
  k_planar<<<nblocks(nv64 * (cout / 8)), kThreads, 0, st>>>(dz, ldz, cout, g.nvox, dzp);
  k_planar<<<nblocks(nv64 * (cin / 8)), kThreads, 0, st>>>(x, ldx, cin, g.nvox, xp);
  const long long len = (long long)cout * 27 * cin;
  const int ci16 = (cin + 15) / 16;
  // about 8192 waves, chunks of whole 32-voxel steps, as many as the partial buffer holds
  const int steps = (g.nvox + 31) / 32;
  long long nch = std::max<long long>(1, std::min<long long>(8192 / (27 * ci16) + 1, steps));
  nch = std::max<long long>(1, std::min<long long>(nch, part_cap / len));
  const int chunk = (int)((steps + nch - 1) / nch) * 32;
  nch = (g.nvox + chunk - 1) / chunk;
  if (nch * len > part_cap) {
    fprintf(stderr, "ls_ws_conv3d_wgrad: partial buffer too small\n");
    return -1;
  }
  const int nwork = (int)(nch * 27 * ci16);
  const unsigned nblk = (unsigned)((nwork + 3) / 4);
  const int cot = (cout + 15) / 16;
#define LS_WS_WGRAD(COT)                                                                                    \
  if (g.Z % 8 == 0)                                                                                         \
    k_wgrad<COT, true><<<nblk, kThreads, 0, st>>>(g, dzp, cout, xp, cin, ci16, chunk, nwork, part);         \
  else                                                                                                      \
    k_wgrad<COT, false><<<nblk, kThreads, 0, st>>>(g, dzp, cout, xp, cin, ci16, chunk, nwork, part);
  if (cot == 1) {
    LS_WS_WGRAD(1)
  } else if (cot == 2) {
    LS_WS_WGRAD(2)
  } else if (cot <= 4) {
    LS_WS_WGRAD(4)
  } else {
    LS_WS_WGRAD(8)
  }
#undef LS_WS_WGRAD
  k_sum_parts<<<nblocks(len), kThreads, 0, st>>>(part, (int)nch, len, dw);
  if (db) {
    const int bch = 4096;
    const int nb = (g.nvox + bch - 1) / bch;
    if ((long long)nb * cout > part_cap) {
      fprintf(stderr, "ls_ws_conv3d_wgrad: partial buffer too small for the bias gradient\n");
      return -1;
    }
    k_bias_parts<<<nb, kThreads, 0, st>>>(dz, ldz, cout, g.nvox, bch, part);
    k_sum_parts<<<nblocks(cout), kThreads, 0, st>>>(part, nb, cout, db);
  }
  return launch_ok("conv3d_wgrad");
}

int ls_ws_pool(const int* dims, const uint16_t* in, int ldi, int c, uint16_t* out, int ldo, void* stream) {
  if (bad_geo(dims) || !in || !out || c % 8 || ldi % 8 || ldo % 8 || dims[1] % 2 || dims[2] % 2 || dims[3] % 2) {
    fprintf(stderr, "ls_ws_pool: invalid arguments\n");
    return -1;
  }
  const Geo g = make_geo(dims);
  k_pool<<<nblocks((long long)g.nvox / 8 * (c / 8)), kThreads, 0, (hipStream_t)stream>>>(g, in, ldi, c, out, ldo);
  return launch_ok("pool");
}

int ls_ws_pool_backward(const int* dims, const uint16_t* dcoarse, int ldc, int c, const uint16_t* s, int lds,
                        const uint16_t* dcat, int ldcat, uint16_t* dz, int ldz, void* stream) {
  if (bad_geo(dims) || !dcoarse || !s || !dz || c % 8 || ldc % 8 || lds % 8 || ldz % 8 || (dcat && ldcat % 8) ||
      dims[1] % 2 || dims[2] % 2 || dims[3] % 2) {
    fprintf(stderr, "ls_ws_pool_backward: invalid arguments\n");
    return -1;
  }
  const Geo g = make_geo(dims);
  k_pool_back<<<nblocks((long long)g.nvox * (c / 8)), kThreads, 0, (hipStream_t)stream>>>(g, dcoarse, ldc, c, s, lds,
                                                                                           dcat, ldcat, dz, ldz);
  return launch_ok("pool_backward");
}

int ls_ws_upsample(const int* dims, const uint16_t* in, int ldi, int c, uint16_t* out, int ldo, void* stream) {
  if (bad_geo(dims) || !in || !out || c % 8 || ldi % 8 || ldo % 8 || dims[1] % 2 || dims[2] % 2 || dims[3] % 2) {
    fprintf(stderr, "ls_ws_upsample: invalid arguments\n");
    return -1;
  }
  const Geo g = make_geo(dims);
  k_upsample<<<nblocks((long long)g.nvox * (c / 8)), kThreads, 0, (hipStream_t)stream>>>(g, in, ldi, c, out, ldo);
  return launch_ok("upsample");
}

int ls_ws_upsample_backward(const int* dims, const uint16_t* dfine, int ldf, int c, uint16_t* dcoarse, int ldc,
                            void* stream) {
  if (bad_geo(dims) || !dfine || !dcoarse || c % 8 || ldf % 8 || ldc % 8 || dims[1] % 2 || dims[2] % 2 ||
      dims[3] % 2) {
    fprintf(stderr, "ls_ws_upsample_backward: invalid arguments\n");
    return -1;
  }
  const Geo g = make_geo(dims);
  k_upsample_back<<<nblocks((long long)g.nvox / 8 * (c / 8)), kThreads, 0, (hipStream_t)stream>>>(g, dfine, ldf, c,
                                                                                                   dcoarse, ldc);
  return launch_ok("upsample_backward");
}

int ls_ws_input(const float* lam, const float* mu, int per_mat, const float* eps_bar, int B, long long nper,
                const float* norm, uint16_t* out, void* stream) {
  if (!lam || !mu || !eps_bar || !norm || !out || per_mat < 1 || B < 1 || nper < 1) {
    fprintf(stderr, "ls_ws_input: invalid arguments\n");
    return -1;
  }
  k_input<<<nblocks((long long)B * nper), kThreads, 0, (hipStream_t)stream>>>(lam, mu, per_mat, eps_bar, B, nper,
                                                                               norm[0], norm[1], norm[2], norm[3], out);
  return launch_ok("input");
}

int ls_ws_target(const float* eps, const float* eps_bar, int B, long long nper, uint16_t* out, void* stream) {
  if (!eps || !eps_bar || !out || B < 1 || nper < 1) return -1;
  k_target<<<nblocks((long long)B * 6 * nper), kThreads, 0, (hipStream_t)stream>>>(eps, eps_bar, B, nper, out);
  return launch_ok("target");
}

int ls_ws_mse(const float* pred, const uint16_t* target, int B, long long nper, float gscale, uint16_t* grad,
              double* part, long long part_cap, double* loss, void* stream) {
  if (!pred || !target || !grad || !part || !loss || B < 1 || nper < 1) return -1;
  const long long nv = (long long)B * nper;
  const unsigned nb = nblocks(nv);
  if ((long long)nb > part_cap) {
    fprintf(stderr, "ls_ws_mse: partial buffer too small\n");
    return -1;
  }
  hipStream_t st = (hipStream_t)stream;
  k_mse<<<nb, kThreads, 0, st>>>(pred, target, B, nper, gscale, grad, part);
  k_sum_f64<<<1, kThreads, 0, st>>>(part, (int)nb, loss);
  return launch_ok("mse");
}


int ls_ws_symgrad(const int* dims, const float* extents, const float* u, int add_unit, float* eps, void* stream) {
  RlGeo g;
  if (!rl_geo(dims, extents, g) || !u || !eps) {
    fprintf(stderr, "ls_ws_symgrad: invalid arguments\n");
    return -1;
  }
  k_symgrad<<<dim3(nblocks(g.nper), g.S), kThreads, 0, (hipStream_t)stream>>>(g, u, add_unit ? 1 : 0, eps);
  return launch_ok("symgrad");
}

int ls_ws_symgrad_backward(const int* dims, const float* extents, const uint16_t* grad_eps, int ldg,
                           uint16_t* grad_u, int ldu, void* stream) {
  RlGeo g;
  if (!rl_geo(dims, extents, g) || !grad_eps || !grad_u || ldg < 36 || ldu < 18 || ldu % 8) {
    fprintf(stderr, "ls_ws_symgrad_backward: invalid arguments\n");
    return -1;
  }
  k_symgrad_back<<<dim3(nblocks(g.nper), g.S), kThreads, 0, (hipStream_t)stream>>>(g, grad_eps, ldg, grad_u, ldu);
  return launch_ok("symgrad_backward");
}

int ls_ws_residual_loss(const int* dims, const float* extents, const float* pred, const uint16_t* target,
                        const float* lam, const float* mu, float wm, float wr, uint16_t* grad, int ldg, float* scratch,
                        long long scratch_cap, double* part, long long part_cap, float* scale, double* loss,
                        void* stream) {
  RlGeo g;
  if (!rl_geo(dims, extents, g) || !pred || !target || !lam || !mu || !grad || !scratch || !part || !scale ||
      !loss || ldg < 36 || ldg % 8) {
    fprintf(stderr, "ls_ws_residual_loss: invalid arguments\n");
    return -1;
  }
  const unsigned nb = nblocks(g.nper);
  if (scratch_cap < (long long)g.S * 54 * g.nper || part_cap < (long long)g.S * nb * 4 * kRlK) {
    fprintf(stderr, "ls_ws_residual_loss: scratch (%lld floats) or partial buffer (%lld) too small\n", scratch_cap,
            part_cap);
    return -1;
  }
  hipStream_t st = (hipStream_t)stream;
  float* sc = scratch;
  float* dv = scratch + (size_t)g.S * 36 * g.nper;
  const dim3 grid(nb, g.S);
  k_rl_stress<<<grid, kThreads, 0, st>>>(g, pred, target, lam, mu, sc, part);
  k_rl_div<<<grid, kThreads, 0, st>>>(g, sc, lam, mu, dv, part);
  k_rl_reduce<<<1, kThreads, 0, st>>>(g.S, (int)nb, wm, wr, part, scale, loss);
  k_rl_grad<<<grid, kThreads, 0, st>>>(g, sc, lam, mu, dv, scale, grad, ldg);
  return launch_ok("residual_loss");
}

int ls_ws_adam(float* w, const float* g, float* m, float* v, long long n, float lr, float b1, float b2, float eps,
               int step, void* stream) {
  if (!w || !g || !m || !v || n < 1 || step < 1) return -1;
  const float c1 = 1.f / (1.f - powf(b1, (float)step));
  const float c2 = 1.f / (1.f - powf(b2, (float)step));
  k_adam<<<nblocks(n), kThreads, 0, (hipStream_t)stream>>>(w, g, m, v, n, lr, b1, b2, eps, c1, c2);
  return launch_ok("adam");
}

int ls_ws_pack(const float* w, int cout, int cin, int cs, uint16_t* wt, int kpad, int rows_t, uint16_t* wd,
               int kpad_d, int rows_d, void* stream) {
  if (!w || !wt || cout < 1 || cin < 1 || kpad < 27 * cin || rows_t < cout || (wd && (kpad_d < 27 * cs || cs < cout ||
                                                                                        rows_d < cin))) {
    fprintf(stderr, "ls_ws_pack: invalid arguments\n");
    return -1;
  }
  const long long n = (long long)rows_t * kpad + (wd ? (long long)rows_d * kpad_d : 0);
  k_pack<<<nblocks(n), kThreads, 0, (hipStream_t)stream>>>(w, cout, cin, cs, wt, kpad, rows_t, wd, kpad_d, rows_d);
  return launch_ok("pack");
}

}  // extern "C"
