// HBM streaming ceilings of the GPU the library runs on: grid-stride float4
// kernels that copy, only read, only write, or read two arrays and write one
// (the read/write mix of the z pass), with non-temporal loads and stores and
// 32768 blocks (the fastest plain streaming form measured on MI355X: a copy
// moves 4.9 TB/s with ordinary accesses and 8192 blocks, 5.2 TB/s so).  The
// benchmark reports the solver's kernels against the published HBM peak and,
// beside it, against these measured ceilings.
#include <hip/hip_runtime.h>

#include <cstdio>

#include "../../include/lippmann_schwinger.h"

namespace {

typedef float f4v_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float4 ldnt4(const float4* p) {
  const f4v_t v = __builtin_nontemporal_load(reinterpret_cast<const f4v_t*>(p));
  return make_float4(v.x, v.y, v.z, v.w);
}

__device__ __forceinline__ void stnt4(float4* p, float4 a) {
  const f4v_t v = {a.x, a.y, a.z, a.w};
  __builtin_nontemporal_store(v, reinterpret_cast<f4v_t*>(p));
}

__global__ void __launch_bounds__(256) k_s_copy(const float4* __restrict__ a, float4* __restrict__ b, long long n) {
  const long long st = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += st) stnt4(&b[i], ldnt4(&a[i]));
}

__global__ void __launch_bounds__(256) k_s_read(const float4* __restrict__ a, float* __restrict__ sink, long long n) {
  const long long st = (long long)gridDim.x * blockDim.x;
  float s = 0.f;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += st) {
    const float4 v = ldnt4(&a[i]);
    s += v.x + v.y + v.z + v.w;
  }
  if (s == -1.f) sink[0] = s;  // keeps the loads; the fill below makes it false
}

__global__ void __launch_bounds__(256) k_s_write(float4* __restrict__ b, long long n, float v) {
  const long long st = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += st)
    stnt4(&b[i], make_float4(v, v, v, v));
}

__global__ void __launch_bounds__(256) k_s_r2w1(const float4* __restrict__ a, const float4* __restrict__ c,
                                               float4* __restrict__ b, long long n) {
  const long long st = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += st) {
    const float4 x = ldnt4(&a[i]), y = ldnt4(&c[i]);
    stnt4(&b[i], make_float4(x.x + y.x, x.y + y.y, x.z + y.z, x.w + y.w));
  }
}

}  // namespace

extern "C" int ls_hbm_stream_probe(long long bytes, int reps, double* gbps, void* stream) {
  if (!gbps || reps < 1 || bytes < (1ll << 20)) return -1;
  const long long n = bytes / 16;
  hipStream_t st = (hipStream_t)stream;
  float4 *a = nullptr, *b = nullptr, *c = nullptr;
  float* sink = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = -1;
  do {
    if (hipMalloc((void**)&a, n * 16) != hipSuccess || hipMalloc((void**)&b, n * 16) != hipSuccess ||
        hipMalloc((void**)&c, n * 16) != hipSuccess || hipMalloc((void**)&sink, 16) != hipSuccess)
      break;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) break;
    const int grid = 32768;
    k_s_write<<<grid, 256, 0, st>>>(a, n, 1.f);
    k_s_write<<<grid, 256, 0, st>>>(c, n, 2.f);
    bool ok = hipGetLastError() == hipSuccess;
    for (int kind = 0; kind < 4 && ok; ++kind) {
      float best = 1e30f;
      for (int r = 0; r < reps && ok; ++r) {
        ok = hipEventRecord(e0, st) == hipSuccess;
        if (kind == 0) k_s_copy<<<grid, 256, 0, st>>>(a, b, n);
        if (kind == 1) k_s_read<<<grid, 256, 0, st>>>(a, sink, n);
        if (kind == 2) k_s_write<<<grid, 256, 0, st>>>(b, n, 3.f);
        if (kind == 3) k_s_r2w1<<<grid, 256, 0, st>>>(a, c, b, n);
        float ms = 0.f;
        ok = ok && hipGetLastError() == hipSuccess && hipEventRecord(e1, st) == hipSuccess &&
             hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
        if (ok && ms < best) best = ms;
      }
      const double moved = (kind == 0 ? 2.0 : kind == 3 ? 3.0 : 1.0) * (double)(n * 16);
      gbps[kind] = moved / (best * 1e-3) / 1e9;
    }
    if (ok) rc = 0;
  } while (0);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  if (c) (void)hipFree(c);
  if (sink) (void)hipFree(sink);
  if (rc != 0) fprintf(stderr, "liblippmannschwinger: ls_hbm_stream_probe failed\n");
  return rc;
}
