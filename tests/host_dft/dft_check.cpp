// Host check of the register DFTs of jaxmaterials_amd/csrc/fft_reg.hpp: every
// compiled length, forward and inverse, against the O(R^2) definition in long
// double on seeded random inputs.  Prints "R forward_err inverse_err" per line
// (relative max-norm errors).  Test infrastructure only.
#include <complex>
#include <cstdio>
#include <random>

#include "../../jaxmaterials_amd/csrc/fft_reg.hpp"

using namespace lsk;

template <int R, bool INV>
double check() {
  std::mt19937 g(R * 2 + INV);
  std::uniform_real_distribution<float> u(-1.f, 1.f);
  float2 v[R];
  std::complex<long double> x[R];
  for (int i = 0; i < R; ++i) {
    v[i] = make_float2(u(g), u(g));
    x[i] = {v[i].x, v[i].y};
  }
  dft_reg<R, INV>(v);
  long double err = 0, nrm = 0;
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int k = 0; k < R; ++k) {
    std::complex<long double> s = 0;
    for (int j = 0; j < R; ++j) s += x[j] * std::polar(1.0L, (INV ? 2 : -2) * pi * j * k / R);
    err = std::max(err, std::abs(s - std::complex<long double>(v[k].x, v[k].y)));
    nrm = std::max(nrm, std::abs(s));
  }
  return (double)(err / nrm);
}

template <int R>
void row() {
  printf("%d %.3e %.3e\n", R, check<R, false>(), check<R, true>());
}

int main() {
  row<2>();
  row<3>();
  row<4>();
  row<5>();
  row<6>();
  row<8>();
  row<12>();
  row<16>();
  row<24>();
  row<32>();
  return 0;
}
