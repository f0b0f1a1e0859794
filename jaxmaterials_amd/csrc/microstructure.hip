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

}  // namespace

extern "C" {

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
