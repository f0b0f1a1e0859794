// Seeded cross-ply fibre microstructures: random sequential adsorption (RSA)
// of circular fibre sections per layer on the host, rasterisation of the Lame
// fields on the GPU.
//
// The rasterisation rule follows LayeredFibresDatasetGenerator.material_properties
// (src/jaxmaterials/data/data.py:267-314): voxel centres x = (i + 1/2) L / n, the
// matrix value everywhere, plus (v_fibre - v_matrix) where
// (c - p0)^2 + (z - p1)^2 < r^2 (strict) with c = y for fibres along x and c = x
// for fibres along y; the in-plane coordinate is periodic (the reference
// appends boundary copies of the fibres, data.py:239-258).  The placement is
// this library's own: the reference places fibres by Yang Chen's movement
// algorithm (libfibres.cc:94-179), the benchmark configurations ask for RSA.
//
// Rules of the placement (layer k of nlayers, thickness t = Lz / nlayers):
//   * fibres of layer k run along x if k is odd, along y if k is even;
//   * centres: in-plane coordinate uniform in [0, Lc), z uniform in
//     [z0 + h, z0 + t - h] with h = min(r, t/2); a voxel belongs to a fibre
//     only if its centre lies in the fibre's layer (fibres thicker than the
//     layer are clipped to it);
//   * a candidate is accepted if its periodic distance to every fibre of the
//     layer is >= 2r; placement stops when n_target = round(vf Lc t / (pi r^2))
//     fibres are placed or after 1000 * (n_target + 1) attempts (jamming);
//   * random numbers: splitmix64 from `seed`, u = (x >> 11) * 2^-53, drawn in
//     the order (layer 0 attempts..., layer 1 attempts, ...), two per attempt.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "../../include/lippmann_schwinger.h"

namespace {

struct SplitMix64 {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

struct Fibre {
  int layer;
  double c;  // in-plane coordinate (y for fibres along x, x for fibres along y)
  double z;
};

// Periodic squared distance; written with explicit temporaries so that the
// host build (-ffp-contract=off) evaluates exactly these IEEE operations.
inline double dist2(double c0, double z0, double c1, double z1, double Lc) {
  double dc = std::fabs(c0 - c1);
  const double wrap = Lc - dc;
  if (wrap < dc) dc = wrap;
  const double dz = z0 - z1;
  const double a = dc * dc;
  const double b = dz * dz;
  return a + b;
}

std::vector<Fibre> place(const int* voxels, const float* extents, int nlayers, double r, double vf, uint64_t seed) {
  std::vector<Fibre> out;
  SplitMix64 rng{seed};
  const double Lx = extents[0], Ly = extents[1], Lz = extents[2];
  const double t = Lz / nlayers;
  const double pi = 3.14159265358979323846;
  for (int k = 0; k < nlayers; ++k) {
    const bool along_x = (k % 2) == 1;
    const double Lc = along_x ? Ly : Lx;
    const double z0 = k * t;
    const double h = std::fmin(r, 0.5 * t);
    const double zlo = z0 + h, zspan = (z0 + t - h) - zlo;
    const double area = vf * Lc * t;
    const long n_target = std::lround(area / (pi * r * r));
    const long max_attempts = 1000 * (n_target + 1);
    const size_t first = out.size();
    long placed = 0;
    for (long a = 0; a < max_attempts && placed < n_target; ++a) {
      const double u1 = rng.uniform();
      const double u2 = rng.uniform();
      const double c = u1 * Lc;
      const double z = zlo + u2 * zspan;
      bool ok = true;
      const double lim = 4.0 * r * r;
      for (size_t i = first; i < out.size(); ++i) {
        if (dist2(c, z, out[i].c, out[i].z, Lc) < lim) {
          ok = false;
          break;
        }
      }
      if (ok) {
        out.push_back({k, c, z});
        ++placed;
      }
    }
  }
  (void)voxels;
  return out;
}

struct DevFibre {
  double c, z;
};

// One thread per voxel; fibres of the voxel's layer only.
__global__ void k_rasterise(int nx, int ny, int nz, double hx, double hy, double hz, double t, int nlayers,
                            const DevFibre* __restrict__ fib, const int* __restrict__ layer_start, double r2,
                            double Lx, double Ly, float lam_m, float mu_m, float lam_f, float mu_f,
                            float* __restrict__ lam, float* __restrict__ mu, unsigned long long* __restrict__ count) {
#pragma clang fp contract(off)
  const long long N = (long long)nx * ny * nz;
  const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  int in_fibre = 0;
  if (v < N) {
    const int iz = (int)(v % nz);
    const int iy = (int)((v / nz) % ny);
    const int ix = (int)(v / ((long long)nz * ny));
    const double x = hx * (0.5 + ix), y = hy * (0.5 + iy), z = hz * (0.5 + iz);
    int k = (int)floor(z / t);
    if (k >= nlayers) k = nlayers - 1;
    const bool along_x = (k % 2) == 1;
    const double c = along_x ? y : x;
    const double Lc = along_x ? Ly : Lx;
    for (int i = layer_start[k]; i < layer_start[k + 1]; ++i) {
      double dc = fabs(c - fib[i].c);
      const double wrap = Lc - dc;
      if (wrap < dc) dc = wrap;
      const double dz = z - fib[i].z;
      const double a = __dmul_rn(dc, dc);
      const double b = __dmul_rn(dz, dz);
      if (__dadd_rn(a, b) < r2) {
        in_fibre = 1;
        break;
      }
    }
    lam[v] = in_fibre ? lam_f : lam_m;
    mu[v] = in_fibre ? mu_f : mu_m;
  }
  // count fibre voxels (one atomic per wave)
  const unsigned long long ballot = __ballot(in_fibre);
  if ((threadIdx.x & 63) == 0 && ballot) atomicAdd(count, (unsigned long long)__popcll(ballot));
}

// ---- layered fibres of the reference generator -----------------------------
//
// data.py:267-314 adds, per channel, (v_fibre - v_matrix) for every fibre
// (boundary copies included) whose disc contains the voxel centre, with the
// test (c - p0)^2 + (Z - p1)^2 < r**2 in float64 (arrays squared by
// multiplication, the NumPy scalar r**2 by libm pow), and (v_void - v_matrix)
// in the void slabs Z < d_void or Z > Lz - d_void.  Every addition is a float64
// sum rounded to the field's dtype (in-place add into a float32 or float64
// array), and all fibre terms are equal, so a voxel's value depends only on
// its void flag and on how many discs contain it.  A fibre of orientation 0
// has c = X, one of orientation 1 has c = Y, so that count is
// cnt0[i][k] + cnt1[j][k]: two small tables of per-line counts, then a fill
// that only writes the fields.

struct LineFibre {
  double p0, p1, r2;
};

// cnt[line][k] = number of fibres (of one orientation) whose disc contains
// (c_line, Z_k), c_line = hc (line + 1/2), Z_k = hz (k + 1/2).
__global__ void __launch_bounds__(256) k_line_counts(int nlines, int nz, double hc, double hz,
                                                     const LineFibre* __restrict__ fib, int nf,
                                                     int* __restrict__ cnt) {
#pragma clang fp contract(off)
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nlines * nz) return;
  const int k = idx % nz, line = idx / nz;
  const double c = hc * (0.5 + line);
  const double z = hz * (0.5 + k);
  int n = 0;
  for (int i = 0; i < nf; ++i) {
    const double dc = c - fib[i].p0;
    const double dz = z - fib[i].p1;
    const double a = dc * dc;
    const double b = dz * dz;
    n += (a + b) < fib[i].r2;
  }
  cnt[idx] = n;
}

struct Channel {
  double matrix, d_fibre, d_void;  // v_matrix, v_fibre - v_matrix, v_void - v_matrix
};

template <typename T>
__device__ __forceinline__ T layered_value(const Channel& ch, bool in_void, int count) {
  T acc = (T)ch.matrix;
  if (in_void) acc = (T)((double)acc + ch.d_void);
  for (int c = 0; c < count; ++c) acc = (T)((double)acc + ch.d_fibre);
  return acc;
}

// out[0] = mu, out[1] = lambda, each [nx][ny][nz]; VEC consecutive z per thread.
template <typename T, int VEC>
__global__ void __launch_bounds__(256) k_fill_layered(int nx, int ny, int nz, const int* __restrict__ cnt0,
                                                      const int* __restrict__ cnt1, double hz, double zlo,
                                                      double zhi, Channel mu, Channel lam, T* __restrict__ out) {
#pragma clang fp contract(off)
  const long long N = (long long)nx * ny * nz;
  const long long v0 = ((long long)blockIdx.x * blockDim.x + threadIdx.x) * VEC;
  if (v0 >= N) return;
  const int k0 = (int)(v0 % nz);
  const long long q = v0 / nz;
  const int j = (int)(q % ny), i = (int)(q / ny);
  T m[VEC], l[VEC];
#pragma unroll
  for (int e = 0; e < VEC; ++e) {
    const int k = k0 + e;
    const double z = hz * (0.5 + k);
    const bool in_void = (z < zlo) || (z > zhi);
    const int count = cnt0[i * nz + k] + cnt1[j * nz + k];
    m[e] = layered_value<T>(mu, in_void, count);
    l[e] = layered_value<T>(lam, in_void, count);
  }
  if constexpr (VEC == 4 && sizeof(T) == 4) {
    *reinterpret_cast<float4*>(out + v0) = make_float4(m[0], m[1], m[2], m[3]);
    *reinterpret_cast<float4*>(out + N + v0) = make_float4(l[0], l[1], l[2], l[3]);
  } else if constexpr (VEC == 2 && sizeof(T) == 8) {
    *reinterpret_cast<double2*>(out + v0) = make_double2(m[0], m[1]);
    *reinterpret_cast<double2*>(out + N + v0) = make_double2(l[0], l[1]);
  } else {
#pragma unroll
    for (int e = 0; e < VEC; ++e) {
      out[v0 + e] = m[e];
      out[N + v0 + e] = l[e];
    }
  }
}

template <typename T, int VEC>
void launch_fill(int nx, int ny, int nz, const int* c0, const int* c1, double hz, double zlo, double zhi,
                 const Channel& mu, const Channel& lam, void* out, hipStream_t st) {
  const long long N = (long long)nx * ny * nz;
  const long long nt = N / VEC;
  const unsigned nb = (unsigned)((nt + 255) / 256);
  k_fill_layered<T, VEC><<<nb, 256, 0, st>>>(nx, ny, nz, c0, c1, hz, zlo, zhi, mu, lam, (T*)out);
}

// libm pow(r, 2) as NumPy evaluates r**2 for a float64 scalar; called through
// a volatile pointer so that the compiler does not replace it by r * r (the
// two differ in the last bit for some r).
double numpy_scalar_square(double r) {
  double (*volatile powp)(double, double) = pow;
  return powp(r, 2.0);
}

}  // namespace

extern "C" {

int ls_rasterise_layered_fibres(const int* voxels, const double* extents, double d_void, int nfibres,
                                const double* fibres, const double* mu3, const double* lambda3, int dtype_bytes,
                                void* d_out, void* stream) {
  if (!voxels || !extents || !mu3 || !lambda3 || !d_out || nfibres < 0 || (nfibres > 0 && !fibres) ||
      (dtype_bytes != 4 && dtype_bytes != 8)) {
    fprintf(stderr, "liblippmannschwinger: ls_rasterise_layered_fibres: invalid arguments\n");
    return -1;
  }
  const int nx = voxels[0], ny = voxels[1], nz = voxels[2];
  if (nx < 1 || ny < 1 || nz < 1 || (long long)nx * nz > (1ll << 30) || (long long)ny * nz > (1ll << 30)) {
    fprintf(stderr, "liblippmannschwinger: ls_rasterise_layered_fibres: invalid grid\n");
    return -1;
  }
  std::vector<LineFibre> lf[2];
  for (int f = 0; f < nfibres; ++f) {
    const double* r = fibres + 4 * f;
    const int o = r[3] != 0.0 ? 1 : 0;
    lf[o].push_back({r[0], r[1], numpy_scalar_square(r[2])});
  }
  hipStream_t st = (hipStream_t)stream;
  LineFibre* d_f = nullptr;
  int* d_cnt = nullptr;
  const size_t n0 = lf[0].size(), n1 = lf[1].size();
  const size_t c0 = (size_t)nx * nz, c1 = (size_t)ny * nz;
  int rc = -1;
  do {
    if (hipMalloc((void**)&d_f, (n0 + n1 + 1) * sizeof(LineFibre)) != hipSuccess) break;
    if (hipMalloc((void**)&d_cnt, (c0 + c1) * sizeof(int)) != hipSuccess) break;
    if (n0 && hipMemcpyAsync(d_f, lf[0].data(), n0 * sizeof(LineFibre), hipMemcpyHostToDevice, st) != hipSuccess)
      break;
    if (n1 && hipMemcpyAsync(d_f + n0, lf[1].data(), n1 * sizeof(LineFibre), hipMemcpyHostToDevice, st) !=
                  hipSuccess)
      break;
    const double hx = extents[0] / nx, hy = extents[1] / ny, hz = extents[2] / nz;
    k_line_counts<<<(unsigned)((c0 + 255) / 256), 256, 0, st>>>(nx, nz, hx, hz, d_f, (int)n0, d_cnt);
    k_line_counts<<<(unsigned)((c1 + 255) / 256), 256, 0, st>>>(ny, nz, hy, hz, d_f + n0, (int)n1, d_cnt + c0);
    const Channel mu{mu3[2], mu3[1] - mu3[2], mu3[0] - mu3[2]};
    const Channel lam{lambda3[2], lambda3[1] - lambda3[2], lambda3[0] - lambda3[2]};
    const double zhi = extents[2] - d_void;
    const bool aligned = ((uintptr_t)d_out % 16) == 0;
    if (dtype_bytes == 4) {
      if (nz % 4 == 0 && aligned)
        launch_fill<float, 4>(nx, ny, nz, d_cnt, d_cnt + c0, hz, d_void, zhi, mu, lam, d_out, st);
      else
        launch_fill<float, 1>(nx, ny, nz, d_cnt, d_cnt + c0, hz, d_void, zhi, mu, lam, d_out, st);
    } else {
      if (nz % 2 == 0 && aligned)
        launch_fill<double, 2>(nx, ny, nz, d_cnt, d_cnt + c0, hz, d_void, zhi, mu, lam, d_out, st);
      else
        launch_fill<double, 1>(nx, ny, nz, d_cnt, d_cnt + c0, hz, d_void, zhi, mu, lam, d_out, st);
    }
    if (hipGetLastError() != hipSuccess) break;
    if (hipStreamSynchronize(st) != hipSuccess) break;
    rc = 0;
  } while (0);
  if (d_f) (void)hipFree(d_f);
  if (d_cnt) (void)hipFree(d_cnt);
  if (rc != 0) fprintf(stderr, "liblippmannschwinger: ls_rasterise_layered_fibres failed\n");
  return rc;
}


int ls_crossply_place(const int* voxels, const float* extents, int nlayers, double radius, double vf,
                      unsigned long long seed, double* out, int cap) {
  if (!voxels || !extents || nlayers < 1 || !(radius > 0.0)) return -1;
  std::vector<Fibre> f = place(voxels, extents, nlayers, radius, vf, seed);
  const int n = (int)f.size();
  for (int i = 0; i < n && i < cap; ++i) {
    out[3 * i] = f[i].layer;
    out[3 * i + 1] = f[i].c;
    out[3 * i + 2] = f[i].z;
  }
  return n;
}

int ls_crossply_lame(const int* voxels, const float* extents, int nlayers, double radius, double vf,
                     unsigned long long seed, double lam_matrix, double mu_matrix, double lam_fibre,
                     double mu_fibre, float* d_lambda, float* d_mu, double* achieved_vf) {
  if (!voxels || !extents || nlayers < 1 || !(radius > 0.0) || !d_lambda || !d_mu) return -1;
  const int nx = voxels[0], ny = voxels[1], nz = voxels[2];
  std::vector<Fibre> f = place(voxels, extents, nlayers, radius, vf, seed);
  std::vector<DevFibre> df(f.size() + 1);
  std::vector<int> start(nlayers + 1, 0);
  for (size_t i = 0; i < f.size(); ++i) {
    df[i] = {f[i].c, f[i].z};
    start[f[i].layer + 1]++;
  }
  for (int k = 0; k < nlayers; ++k) start[k + 1] += start[k];
  DevFibre* d_f = nullptr;
  int* d_start = nullptr;
  unsigned long long* d_count = nullptr;
  int rc = (int)f.size();
  if (hipMalloc(&d_f, df.size() * sizeof(DevFibre)) != hipSuccess ||
      hipMalloc(&d_start, start.size() * sizeof(int)) != hipSuccess ||
      hipMalloc(&d_count, sizeof(unsigned long long)) != hipSuccess ||
      hipMemcpy(d_f, df.data(), df.size() * sizeof(DevFibre), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d_start, start.data(), start.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(d_count, 0, sizeof(unsigned long long)) != hipSuccess) {
    fprintf(stderr, "liblippmannschwinger: rasteriser allocation failed\n");
    rc = -1;
  } else {
    const long long N = (long long)nx * ny * nz;
    const int bs = 256;
    const long long nb = (N + bs - 1) / bs;
    k_rasterise<<<(unsigned)nb, bs>>>(nx, ny, nz, (double)extents[0] / nx, (double)extents[1] / ny,
                                      (double)extents[2] / nz, (double)extents[2] / nlayers, nlayers, d_f, d_start,
                                      radius * radius, extents[0], extents[1], (float)lam_matrix, (float)mu_matrix,
                                      (float)lam_fibre, (float)mu_fibre, d_lambda, d_mu, d_count);
    unsigned long long cnt = 0;
    if (hipGetLastError() != hipSuccess ||
        hipMemcpy(&cnt, d_count, sizeof(cnt), hipMemcpyDeviceToHost) != hipSuccess) {
      fprintf(stderr, "liblippmannschwinger: rasteriser launch failed\n");
      rc = -1;
    }
    if (achieved_vf) *achieved_vf = (double)cnt / (double)N;
  }
  hipFree(d_f);
  hipFree(d_start);
  hipFree(d_count);
  return rc;
}

}  // extern "C"
