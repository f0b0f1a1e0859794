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
// The engine is templated on the complex type: float2 for the iteration,
// double2 for the rare FP64 base evaluations (tw64: FP64 table).
//
// Forward is unnormalised with exp(-i ...), inverse is unnormalised with
// exp(+i ...), matching the FFT conventions of the reference
// (src/jaxmaterials/solver/lippmann_schwinger.py:192, 216; cuFFT R2C/C2R at
// cuda/src/lippmann_schwinger.cu:119-120).
#pragma once
#include <hip/hip_runtime.h>

namespace lsk {

constexpr int kMaxStages = 16;
constexpr int kThreads = 256;  // default block size (4 waves of 64)

// Division by a runtime-invariant divisor d >= 1 for numerators 0 <= n < 2^31:
// q = (umulhi(n, m) + n) >> s with s = ceil(log2 d),
// m = floor(2^32 (2^s - d) / d) + 1 (Granlund & Montgomery 1994).  Integer
// division by a runtime value costs tens of instructions on the GPU; the FFT
// index arithmetic uses these instead.
struct FastDiv {
  unsigned d, m, s;
};

inline FastDiv make_fastdiv(unsigned d) {
  FastDiv f;
  f.d = d;
  unsigned s = 0;
  while ((1ull << s) < d) ++s;
  f.s = s;
  f.m = (unsigned)((((1ull << 32) * ((1ull << s) - d)) / d) + 1);
  return f;
}

__device__ __forceinline__ int fdiv(int n, const FastDiv& f) {
  return (int)((__umulhi((unsigned)n, f.m) + (unsigned)n) >> f.s);
}

struct AxisPlan {
  int n;
  int nstages;  // 0 => direct DFT
  int radix[kMaxStages];
  FastDiv f_ns[kMaxStages];  // Ns = product of the earlier radices
  FastDiv f_nr[kMaxStages];  // n / radix
  FastDiv f_n;
  const float2* tw;    // n entries, exp(-2 pi i t / n)
  const double2* tw64;  // the same in FP64
};

// LDS index of element x of line l: lines of stride ls with element stride es,
// or (PAD) contiguous lines with one pad slot per 16 elements, which spreads the
// strided Stockham accesses over the LDS banks.
template <bool PAD>
__device__ __forceinline__ int lds_idx(int l, int x, int es, int ls) {
  return PAD ? l * ls + x + (x >> 4) : l * ls + x * es;
}

template <typename V>
struct CT;
template <>
struct CT<float2> {
  using S = float;
  static __device__ __forceinline__ const float2* tw(const AxisPlan& p) { return p.tw; }
  static __device__ __forceinline__ float2 mk(float x, float y) { return make_float2(x, y); }
};
template <>
struct CT<double2> {
  using S = double;
  static __device__ __forceinline__ const double2* tw(const AxisPlan& p) { return p.tw64; }
  static __device__ __forceinline__ double2 mk(double x, double y) { return make_double2(x, y); }
};

template <typename V>
__device__ __forceinline__ V cadd(V a, V b) { return CT<V>::mk(a.x + b.x, a.y + b.y); }
template <typename V>
__device__ __forceinline__ V csub(V a, V b) { return CT<V>::mk(a.x - b.x, a.y - b.y); }
template <typename V>
__device__ __forceinline__ V cmul(V a, V b) { return CT<V>::mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
// a * conj(b)
template <typename V>
__device__ __forceinline__ V cmulc(V a, V b) { return CT<V>::mk(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y); }
template <typename V>
__device__ __forceinline__ V cscale(V a, typename CT<V>::S s) { return CT<V>::mk(a.x * s, a.y * s); }
// multiply by -i (forward) or +i (inverse)
template <bool INV, typename V>
__device__ __forceinline__ V mul_mi(V a) {
  return INV ? CT<V>::mk(-a.y, a.x) : CT<V>::mk(a.y, -a.x);
}

template <bool INV, typename V>
__device__ __forceinline__ void dft2(V* v) {
  V a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <bool INV, typename V>
__device__ __forceinline__ void dft4(V* v) {
  V t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
  V t2 = cadd(v[1], v[3]), t3 = mul_mi<INV>(csub(v[1], v[3]));
  v[0] = cadd(t0, t2);
  v[2] = csub(t0, t2);
  v[1] = cadd(t1, t3);
  v[3] = csub(t1, t3);
}

template <bool INV, typename V>
__device__ __forceinline__ void dft8(V* v) {
  using S = typename CT<V>::S;
  constexpr S s = S(0.70710678118654752440);
  V e[4] = {v[0], v[2], v[4], v[6]};
  V o[4] = {v[1], v[3], v[5], v[7]};
  dft4<INV>(e);
  dft4<INV>(o);
  // W8^k for k = 1, 2, 3 (conjugated for the inverse)
  V o1 = INV ? CT<V>::mk(s * (o[1].x - o[1].y), s * (o[1].x + o[1].y))
             : CT<V>::mk(s * (o[1].x + o[1].y), s * (o[1].y - o[1].x));
  V o2 = mul_mi<INV>(o[2]);
  V o3 = INV ? CT<V>::mk(-s * (o[3].x + o[3].y), s * (o[3].x - o[3].y))
             : CT<V>::mk(s * (o[3].y - o[3].x), -s * (o[3].x + o[3].y));
  v[0] = cadd(e[0], o[0]);
  v[4] = csub(e[0], o[0]);
  v[1] = cadd(e[1], o1);
  v[5] = csub(e[1], o1);
  v[2] = cadd(e[2], o2);
  v[6] = csub(e[2], o2);
  v[3] = cadd(e[3], o3);
  v[7] = csub(e[3], o3);
}

template <bool INV, typename V>
__device__ __forceinline__ void dft3(V* v) {
  using S = typename CT<V>::S;
  constexpr S s3 = S(0.86602540378443864676);
  V t1 = cadd(v[1], v[2]), t2 = csub(v[1], v[2]);
  V y0 = cadd(v[0], t1);
  V m = CT<V>::mk(v[0].x - S(0.5) * t1.x, v[0].y - S(0.5) * t1.y);
  V it2 = mul_mi<INV>(cscale(t2, s3));  // -i s3 t2 (forward)
  v[0] = y0;
  v[1] = cadd(m, it2);
  v[2] = csub(m, it2);
}

template <bool INV, typename V>
__device__ __forceinline__ void dft5(V* v) {
  using S = typename CT<V>::S;
  constexpr S c1 = S(0.30901699437494742410);   // cos(2 pi / 5)
  constexpr S c2 = S(-0.80901699437494742410);  // cos(4 pi / 5)
  constexpr S s1 = S(0.95105651629515357212);   // sin(2 pi / 5)
  constexpr S s2 = S(0.58778525229247312917);   // sin(4 pi / 5)
  V t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
  V t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
  V a1 = CT<V>::mk(v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y);
  V a2 = CT<V>::mk(v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y);
  V b1 = CT<V>::mk(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y);
  V b2 = CT<V>::mk(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y);
  V ib1 = mul_mi<INV>(b1), ib2 = mul_mi<INV>(b2);
  v[0] = cadd(v[0], cadd(t1, t2));
  v[1] = cadd(a1, ib1);
  v[4] = csub(a1, ib1);
  v[2] = cadd(a2, ib2);
  v[3] = csub(a2, ib2);
}

template <int R, bool INV, typename V>
__device__ __forceinline__ void dftR(V* v) {
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
__device__ __forceinline__ void split_item(int w, const FastDiv& f_lines, const FastDiv& f_per_line, int& l,
                                           int& j) {
  if (LINE_FAST) {
    j = fdiv(w, f_lines);
    l = w - j * (int)f_lines.d;
  } else {
    l = fdiv(w, f_per_line);
    j = w - l * (int)f_per_line.d;
  }
}

// One Stockham stage of radix R, in place in LDS, by a block of NT threads.
// E = register elements per thread; the host guarantees
// ceil(nlines * n / R / NT) * R <= E.
template <int R, bool INV, bool LINE_FAST, bool PAD, int E, int NT, typename V>
__device__ __forceinline__ void stockham_stage(V* buf, int n, const FastDiv& f_ns, const FastDiv& f_nr,
                                               const FastDiv& f_lines, int nlines, int es, int ls,
                                               const V* __restrict__ tw) {
  constexpr int NB = E / R;
  const int tid = threadIdx.x;
  const int nR = (int)f_nr.d;
  const int Ns = (int)f_ns.d;
  const int total = nlines * nR;
  const int tws = n / (Ns * R);
  V v[NB * R];
  int dst0[NB];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int w = tid + b * NT;
    dst0[b] = -1;
    if (w < total) {
      int l, j;
      split_item<LINE_FAST>(w, f_lines, f_nr, l, j);
      const int jq = fdiv(j, f_ns);
      const int k = j - jq * Ns;
#pragma unroll
      for (int q = 0; q < R; ++q) {
        V x = buf[lds_idx<PAD>(l, j + q * nR, es, ls)];
        if (q > 0) {
          V t = tw[k * q * tws];
          x = INV ? cmulc(x, t) : cmul(x, t);
        }
        v[b * R + q] = x;
      }
      dftR<R, INV>(&v[b * R]);
      dst0[b] = l * 0x10000 + (jq * Ns * R + k);  // (line, first output index), both < 2^16
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    if (dst0[b] >= 0) {
      const int l = dst0[b] >> 16, d = dst0[b] & 0xffff;
#pragma unroll
      for (int q = 0; q < R; ++q) buf[lds_idx<PAD>(l, d + q * Ns, es, ls)] = v[b * R + q];
    }
  }
  __syncthreads();
}

// Direct DFT of every line (any n); requires ceil(nlines * n / NT) <= E.
template <bool INV, bool LINE_FAST, bool PAD, int E, int NT, typename V>
__device__ __forceinline__ void direct_dft(V* buf, const FastDiv& f_n, const FastDiv& f_lines, int nlines, int es,
                                           int ls, const V* __restrict__ tw) {
  const int tid = threadIdx.x;
  const int n = (int)f_n.d;
  const int total = nlines * n;
  V v[E];
#pragma unroll
  for (int b = 0; b < E; ++b) {
    const int w = tid + b * NT;
    if (w < total) {
      int l, k;
      split_item<LINE_FAST>(w, f_lines, f_n, l, k);
      V acc = CT<V>::mk(0, 0);
      int idx = 0;
      for (int j = 0; j < n; ++j) {
        V x = buf[lds_idx<PAD>(l, j, es, ls)];
        V t = tw[idx];
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
    const int w = tid + b * NT;
    if (w < total) {
      int l, k;
      split_item<LINE_FAST>(w, f_lines, f_n, l, k);
      buf[lds_idx<PAD>(l, k, es, ls)] = v[b];
    }
  }
  __syncthreads();
}

// One Stockham stage of a radix the butterflies above do not cover (a factor
// of the length with primes above 5, e.g. 7, 11, 13 or a product of them):
// every work item computes one output of an R-point DFT of a strided
// subsequence, with one table twiddle per term.  Output q_out of the
// butterfly (jq, k), j = jq Ns + k, is
//   sum_q x[j + q nR] w^(q (k + q_out Ns) tws),  w = exp(-+2 pi i / n),
// tws = n / (Ns R) (the stage twiddle w^(q k tws) and the R-point root
// w^(q q_out n / R) combined), written to jq Ns R + q_out Ns + k, the place
// stockham_stage writes it.  Requires ceil(nlines * n / NT) <= E.
// f_nsR divides by Ns R (the next stage's Ns).
template <bool INV, bool LINE_FAST, bool PAD, int E, int NT, typename V>
__device__ __forceinline__ void generic_stage(V* buf, int n, int R, const FastDiv& f_ns, const FastDiv& f_nsR,
                                              const FastDiv& f_n, const FastDiv& f_lines, int nlines, int es,
                                              int ls, const V* __restrict__ tw) {
  const int tid = threadIdx.x;
  const int Ns = (int)f_ns.d;
  const int NsR = (int)f_nsR.d;
  const int nR = n / R;
  const int tws = n / NsR;
  const int total = nlines * n;
  V v[E];
#pragma unroll
  for (int b = 0; b < E; ++b) {
    const int w = tid + b * NT;
    if (w < total) {
      int l, o;
      split_item<LINE_FAST>(w, f_lines, f_n, l, o);
      const int jq = fdiv(o, f_nsR);
      const int rem = o - jq * NsR;
      const int qo = fdiv(rem, f_ns);
      const int k = rem - qo * Ns;
      const int j = jq * Ns + k;
      const int step = (k + qo * Ns) * tws;  // < n
      V acc = CT<V>::mk(0, 0);
      int idx = 0;
#pragma unroll 1
      for (int q = 0; q < R; ++q) {
        V x = buf[lds_idx<PAD>(l, j + q * nR, es, ls)];
        V tq = tw[idx];
        acc = cadd(acc, INV ? cmulc(x, tq) : cmul(x, tq));
        idx += step;
        if (idx >= n) idx -= n;
      }
      v[b] = acc;
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < E; ++b) {
    const int w = tid + b * NT;
    if (w < total) {
      int l, o;
      split_item<LINE_FAST>(w, f_lines, f_n, l, o);
      buf[lds_idx<PAD>(l, o, es, ls)] = v[b];
    }
  }
  __syncthreads();
}

// Transform all lines in place.  Must be called by every thread of the block.
// f_lines (LINE_FAST layouts) divides by nlines; lines must be < 2^15 and
// line lengths < 2^16.  GEN = false leaves generic_stage out (it raises the
// register count of the kernel even where it never runs); only for plans
// without such a stage (the z axis, whose lengths with primes above 5 take
// the direct DFT).
template <bool INV, bool LINE_FAST, bool PAD, int E, int NT, bool GEN = true, typename V>
__device__ __forceinline__ void fft_lines(V* buf, const AxisPlan& p, const FastDiv& f_lines, int nlines, int es,
                                          int ls) {
  const V* tw = CT<V>::tw(p);
  if (p.nstages == 0) {
    direct_dft<INV, LINE_FAST, PAD, E, NT>(buf, p.f_n, f_lines, nlines, es, ls, tw);
    return;
  }
  for (int s = 0; s < p.nstages; ++s) {
    const int R = p.radix[s];
#define LS_STAGE(RV) \
  stockham_stage<RV, INV, LINE_FAST, PAD, E, NT>(buf, p.n, p.f_ns[s], p.f_nr[s], f_lines, nlines, es, ls, tw)
    switch (R) {
      case 8: LS_STAGE(8); break;
      case 4: LS_STAGE(4); break;
      case 2: LS_STAGE(2); break;
      case 3: LS_STAGE(3); break;
      case 5: LS_STAGE(5); break;
      default:
        if constexpr (GEN)
          generic_stage<INV, LINE_FAST, PAD, E, NT>(buf, p.n, R, p.f_ns[s], p.f_ns[s + 1], p.f_n, f_lines, nlines,
                                                    es, ls, tw);
        break;
    }
#undef LS_STAGE
  }
}

}  // namespace lsk
