// Batched complex FFTs of lines held in LDS, for gfx950 (CDNA4, wave64).
//
// A workgroup holds `nlines` lines of length n in LDS; element x of line l sits
// at buf[l * ls + x * es].  The transform is a mixed-radix Stockham autosort
// (radices 8, 4, 2, 3, 5) with twiddles from a per-axis FP32 table computed in
// FP64 on the host: tw[t] = exp(-2 pi i t / n).  Each stage stages every
// thread's butterfly inputs in registers, then barriers, then writes, so one
// LDS buffer suffices (in-place).  Lengths with a prime factor above 5 use a
// direct O(n^2) DFT per line (the reference accepts any grid size; its tests
// use 37, 43, 47, 61 and 57 = 3 * 19).
//
// Forward is unnormalised with exp(-i ...), inverse is unnormalised with
// exp(+i ...), matching the FFT conventions of the reference
// (src/jaxmaterials/solver/lippmann_schwinger.py:192, 216; cuFFT R2C/C2R at
// cuda/src/lippmann_schwinger.cu:119-120).
#pragma once
#include <hip/hip_runtime.h>

namespace lsk {

constexpr int kMaxStages = 16;
constexpr int kThreads = 256;  // every FFT kernel runs 4 waves of 64

struct AxisPlan {
  int n;
  int nstages;  // 0 => direct DFT
  int radix[kMaxStages];
  const float2* tw;  // n entries, exp(-2 pi i t / n)
};

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// multiply by -i (forward) or +i (inverse)
template <bool INV>
__device__ __forceinline__ float2 mul_mi(float2 a) {
  return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

template <bool INV>
__device__ __forceinline__ void dft2(float2* v) {
  float2 a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <bool INV>
__device__ __forceinline__ void dft4(float2* v) {
  float2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
  float2 t2 = cadd(v[1], v[3]), t3 = mul_mi<INV>(csub(v[1], v[3]));
  v[0] = cadd(t0, t2);
  v[2] = csub(t0, t2);
  v[1] = cadd(t1, t3);
  v[3] = csub(t1, t3);
}

template <bool INV>
__device__ __forceinline__ void dft8(float2* v) {
  constexpr float s = 0.70710678118654752440f;
  float2 e[4] = {v[0], v[2], v[4], v[6]};
  float2 o[4] = {v[1], v[3], v[5], v[7]};
  dft4<INV>(e);
  dft4<INV>(o);
  // W8^k for k = 1, 2, 3 (conjugated for the inverse)
  float2 o1 = INV ? make_float2(s * (o[1].x - o[1].y), s * (o[1].x + o[1].y))
                  : make_float2(s * (o[1].x + o[1].y), s * (o[1].y - o[1].x));
  float2 o2 = mul_mi<INV>(o[2]);
  float2 o3 = INV ? make_float2(-s * (o[3].x + o[3].y), s * (o[3].x - o[3].y))
                  : make_float2(s * (o[3].y - o[3].x), -s * (o[3].x + o[3].y));
  v[0] = cadd(e[0], o[0]);
  v[4] = csub(e[0], o[0]);
  v[1] = cadd(e[1], o1);
  v[5] = csub(e[1], o1);
  v[2] = cadd(e[2], o2);
  v[6] = csub(e[2], o2);
  v[3] = cadd(e[3], o3);
  v[7] = csub(e[3], o3);
}

template <bool INV>
__device__ __forceinline__ void dft3(float2* v) {
  constexpr float s3 = 0.86602540378443864676f;
  float2 t1 = cadd(v[1], v[2]), t2 = csub(v[1], v[2]);
  float2 y0 = cadd(v[0], t1);
  float2 m = make_float2(v[0].x - 0.5f * t1.x, v[0].y - 0.5f * t1.y);
  float2 it2 = mul_mi<INV>(cscale(t2, s3));  // -i s3 t2 (forward)
  v[0] = y0;
  v[1] = cadd(m, it2);
  v[2] = csub(m, it2);
}

template <bool INV>
__device__ __forceinline__ void dft5(float2* v) {
  constexpr float c1 = 0.30901699437494742410f;   // cos(2 pi / 5)
  constexpr float c2 = -0.80901699437494742410f;  // cos(4 pi / 5)
  constexpr float s1 = 0.95105651629515357212f;   // sin(2 pi / 5)
  constexpr float s2 = 0.58778525229247312917f;   // sin(4 pi / 5)
  float2 t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
  float2 t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
  float2 a1 = make_float2(v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y);
  float2 a2 = make_float2(v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y);
  float2 b1 = make_float2(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y);
  float2 b2 = make_float2(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y);
  float2 ib1 = mul_mi<INV>(b1), ib2 = mul_mi<INV>(b2);
  v[0] = cadd(v[0], cadd(t1, t2));
  v[1] = cadd(a1, ib1);
  v[4] = csub(a1, ib1);
  v[2] = cadd(a2, ib2);
  v[3] = csub(a2, ib2);
}

template <int R, bool INV>
__device__ __forceinline__ void dftR(float2* v) {
  if constexpr (R == 2) dft2<INV>(v);
  else if constexpr (R == 3) dft3<INV>(v);
  else if constexpr (R == 4) dft4<INV>(v);
  else if constexpr (R == 5) dft5<INV>(v);
  else dft8<INV>(v);
}

// Map a work item to (line, position).  With LINE_FAST consecutive work items
// walk across lines (use when lines are adjacent in LDS, i.e. ls == 1), else
// along a line (use when elements are adjacent, es == 1).
template <bool LINE_FAST>
__device__ __forceinline__ void split_item(int w, int nlines, int per_line, int& l, int& j) {
  if (LINE_FAST) {
    j = w / nlines;
    l = w - j * nlines;
  } else {
    l = w / per_line;
    j = w - l * per_line;
  }
}

// One Stockham stage of radix R, in place in LDS.  E = register elements per
// thread; the host guarantees ceil(nlines * n / R / kThreads) * R <= E.
template <int R, bool INV, bool LINE_FAST, int E>
__device__ __forceinline__ void stockham_stage(float2* buf, int n, int Ns, int nlines, int es, int ls,
                                               const float2* __restrict__ tw) {
  constexpr int NB = E / R;
  const int tid = threadIdx.x;
  const int nR = n / R;
  const int total = nlines * nR;
  const int tws = n / (Ns * R);
  float2 v[NB * R];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int w = tid + b * kThreads;
    if (w < total) {
      int l, j;
      split_item<LINE_FAST>(w, nlines, nR, l, j);
      const float2* src = buf + l * ls;
      const int k = j % Ns;
#pragma unroll
      for (int q = 0; q < R; ++q) {
        float2 x = src[(j + q * nR) * es];
        if (q > 0) {
          float2 t = tw[k * q * tws];
          x = INV ? cmulc(x, t) : cmul(x, t);
        }
        v[b * R + q] = x;
      }
      dftR<R, INV>(&v[b * R]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int w = tid + b * kThreads;
    if (w < total) {
      int l, j;
      split_item<LINE_FAST>(w, nlines, nR, l, j);
      float2* dst = buf + l * ls;
      const int k = j % Ns;
      const int d = (j / Ns) * Ns * R + k;
#pragma unroll
      for (int q = 0; q < R; ++q) dst[(d + q * Ns) * es] = v[b * R + q];
    }
  }
  __syncthreads();
}

// Direct DFT of every line (any n); requires ceil(nlines * n / kThreads) <= E.
template <bool INV, bool LINE_FAST, int E>
__device__ __forceinline__ void direct_dft(float2* buf, int n, int nlines, int es, int ls,
                                           const float2* __restrict__ tw) {
  const int tid = threadIdx.x;
  const int total = nlines * n;
  float2 v[E];
#pragma unroll
  for (int b = 0; b < E; ++b) {
    const int w = tid + b * kThreads;
    if (w < total) {
      int l, k;
      split_item<LINE_FAST>(w, nlines, n, l, k);
      const float2* src = buf + l * ls;
      float2 acc = make_float2(0.f, 0.f);
      int idx = 0;
      for (int j = 0; j < n; ++j) {
        float2 x = src[j * es];
        float2 t = tw[idx];
        acc = cadd(acc, INV ? cmulc(x, t) : cmul(x, t));
        idx += k;
        if (idx >= n) idx -= n;
      }
      v[b] = acc;
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < E; ++b) {
    const int w = tid + b * kThreads;
    if (w < total) {
      int l, k;
      split_item<LINE_FAST>(w, nlines, n, l, k);
      buf[l * ls + k * es] = v[b];
    }
  }
  __syncthreads();
}

// Transform all lines in place.  Must be called by every thread of the block.
template <bool INV, bool LINE_FAST, int E>
__device__ void fft_lines(float2* buf, const AxisPlan& p, int nlines, int es, int ls) {
  if (p.nstages == 0) {
    direct_dft<INV, LINE_FAST, E>(buf, p.n, nlines, es, ls, p.tw);
    return;
  }
  int Ns = 1;
  for (int s = 0; s < p.nstages; ++s) {
    const int R = p.radix[s];
    switch (R) {
      case 8: stockham_stage<8, INV, LINE_FAST, E>(buf, p.n, Ns, nlines, es, ls, p.tw); break;
      case 4: stockham_stage<4, INV, LINE_FAST, E>(buf, p.n, Ns, nlines, es, ls, p.tw); break;
      case 2: stockham_stage<2, INV, LINE_FAST, E>(buf, p.n, Ns, nlines, es, ls, p.tw); break;
      case 3: stockham_stage<3, INV, LINE_FAST, E>(buf, p.n, Ns, nlines, es, ls, p.tw); break;
      default: stockham_stage<5, INV, LINE_FAST, E>(buf, p.n, Ns, nlines, es, ls, p.tw); break;
    }
    Ns *= R;
  }
}

}  // namespace lsk
