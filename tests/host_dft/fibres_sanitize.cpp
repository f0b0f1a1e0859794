// Host driver of the library's fibre movement (jaxmaterials_amd/csrc/fibres.cpp)
// for tests/test_host_sanitizers.py: runs dist_periodic, initialise_rng and
// move_fibres on several seeded layouts (odd and even fibre counts, a layout
// that needs many sweeps, a one-fibre box) so AddressSanitizer and
// UndefinedBehaviorSanitizer see every code path.  Prints one checksum line
// per case.  Test infrastructure only.
#include <cstdio>
#include <random>
#include <vector>

#include "../../include/fibres.h"

static double run_case(unsigned n, double L0, double L1, double r, unsigned repeats, unsigned seed) {
  std::mt19937 g(seed);
  std::uniform_real_distribution<double> u(0.0, 1.0);
  std::vector<double> loc(2 * n), rad(n, r);
  for (unsigned i = 0; i < n; ++i) {
    loc[2 * i] = u(g) * L0;
    loc[2 * i + 1] = u(g) * L1;
    rad[i] = r * (0.8 + 0.4 * u(g));
  }
  const double dom[2] = {L0, L1};
  std::vector<char> state(10000);
  unsigned len = 0;
  initialise_rng(735283u + seed, &len, state.data());
  move_fibres(n, dom, rad.data(), loc.data(), repeats, 1e-3, 5000, &len, state.data());
  std::vector<double> dist(n);
  double sum = 0.0;
  for (unsigned i = 0; i < n; ++i) {
    dist_periodic(n, &loc[2 * i], loc.data(), dom, dist.data());
    for (unsigned j = 0; j < n; ++j) sum += dist[j];
  }
  return sum;
}

int main() {
  std::printf("%.17g\n", run_case(1, 4.0, 3.0, 0.5, 3, 1));
  std::printf("%.17g\n", run_case(7, 10.0, 6.0, 0.9, 5, 2));
  std::printf("%.17g\n", run_case(16, 12.0, 12.0, 1.0, 10, 3));
  std::printf("%.17g\n", run_case(41, 20.0, 16.0, 1.0, 2, 4));
  return 0;
}
