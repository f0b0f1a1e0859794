// Register-resident DFTs of compile-time length for gfx950.
//
// dft_reg<R, INV>(v) transforms R complex values held in registers in place
// (natural order in and out); forward uses exp(-2 pi i jk / R), inverse
// exp(+2 pi i jk / R), both unnormalised.  Lengths 2, 3, 4, 5, 8 are the
// butterflies of fft_lds.hpp; 6, 12, 16, 24 and 32 are composed from them by
// one Cooley-Tukey split R = R1 * R2 whose twiddles are compile-time
// constants, so a whole transform is straight-line code with no table loads.
#pragma once
#include "fft_lds.hpp"

namespace lsk {

// cos(2 pi k / 96) for any integer k: the first quadrant to full double
// precision, the rest by symmetry.
constexpr double cos96(int k) {
  const double q[25] = {1.0,
                        0.997858923238603506738,
                        0.991444861373810411145,
                        0.980785280403230449126,
                        0.96592582628906828675,
                        0.946930129495105664256,
                        0.923879532511286756128,
                        0.896872741532688303894,
                        0.866025403784438646764,
                        0.831469612302545237079,
                        0.79335334029123516458,
                        0.751839807478977396408,
                        0.707106781186547524401,
                        0.659345815100068868425,
                        0.608761429008720639416,
                        0.555570233019602224743,
                        0.5,
                        0.442288690219001281995,
                        0.382683432365089771728,
                        0.321439465303161580701,
                        0.258819045102520762349,
                        0.195090322016128267848,
                        0.130526192220051591548,
                        0.0654031292301430668153,
                        0.0};
  k %= 96;
  if (k < 0) k += 96;
  if (k <= 24) return q[k];
  if (k <= 48) return -q[48 - k];
  if (k <= 72) return -q[k - 48];
  return q[96 - k];
}

// x * W_R^e with W_R = exp(-2 pi i / R) (conjugated for the inverse); R must
// divide 96.  Multiples of a quarter turn are exact swaps / negations.
template <int R, int E, bool INV, typename V>
__device__ __forceinline__ V twc(V x) {
  static_assert(96 % R == 0, "twiddle table resolution is 2 pi / 96");
  using S = typename CT<V>::S;
  constexpr int e = ((E % R) + R) % R;
  if constexpr (e == 0) {
    return x;
  } else if constexpr (4 * e == R) {
    return mul_mi<INV>(x);  // -i (forward)
  } else if constexpr (2 * e == R) {
    return CT<V>::mk(-x.x, -x.y);
  } else if constexpr (4 * e == 3 * R) {
    return mul_mi<!INV>(x);  // +i (forward)
  } else {
    constexpr int k = e * (96 / R);
    constexpr S c = S(cos96(k));
    constexpr S s = S(cos96(k - 24));  // sin(2 pi e / R)
    // forward: (c - i s) x ; inverse: (c + i s) x
    if constexpr (INV)
      return CT<V>::mk(c * x.x - s * x.y, c * x.y + s * x.x);
    else
      return CT<V>::mk(c * x.x + s * x.y, c * x.y - s * x.x);
  }
}

template <int R, bool INV, typename V>
__device__ __forceinline__ void dft_reg(V* v);

// w[t R1 + k1] *= W_R^(t k1) for every (t, k1), each exponent a constant.
template <int R, int R1, int I, bool INV, typename V>
__device__ __forceinline__ void tw_all(V* w) {
  if constexpr (I < R) {
    w[I] = twc<R, (I / R1) * (I % R1), INV>(w[I]);
    tw_all<R, R1, I + 1, INV>(w);
  }
}

// R = R1 * R2: x index = t + R2 j (t < R2, j < R1); R1-point DFTs over j,
// twiddle W_R^(t k1), R2-point DFTs over t; output index k1 + R1 k2.
template <int R1, int R2, bool INV, typename V>
__device__ __forceinline__ void dft_split(V* v) {
  constexpr int R = R1 * R2;
  V w[R];
#pragma unroll
  for (int t = 0; t < R2; ++t) {
    V u[R1];
#pragma unroll
    for (int j = 0; j < R1; ++j) u[j] = v[t + R2 * j];
    dft_reg<R1, INV>(u);
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) w[t * R1 + k1] = u[k1];
  }
  tw_all<R, R1, 0, INV>(w);
#pragma unroll
  for (int k1 = 0; k1 < R1; ++k1) {
    V u[R2];
#pragma unroll
    for (int t = 0; t < R2; ++t) u[t] = w[t * R1 + k1];
    dft_reg<R2, INV>(u);
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) v[k1 + R1 * k2] = u[k2];
  }
}

template <int R, bool INV, typename V>
__device__ __forceinline__ void dft_reg(V* v) {
  if constexpr (R == 1) {
  } else if constexpr (R == 2 || R == 3 || R == 4 || R == 5 || R == 8) {
    dftR<R, INV>(v);
  } else if constexpr (R == 6) {
    dft_split<2, 3, INV>(v);
  } else if constexpr (R == 12) {
    dft_split<4, 3, INV>(v);
  } else if constexpr (R == 16) {
    dft_split<4, 4, INV>(v);
  } else if constexpr (R == 24) {
    dft_split<8, 3, INV>(v);
  } else if constexpr (R == 32) {
    dft_split<8, 4, INV>(v);
  } else {
    static_assert(R == 2, "unsupported register DFT length");
  }
}

}  // namespace lsk
