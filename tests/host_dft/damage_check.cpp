// Host build of the damage split (jaxmaterials_amd/csrc/damage.hpp) for
// tests/test_host_damage.py: reads lines "lam mu v0 .. v5" from stdin and
// prints sigma+ (six values) and psi+ per line.  Test infrastructure only.
#include <cstdio>

#include "../../jaxmaterials_amd/csrc/damage.hpp"

int main() {
  double lam, mu, v[6];
  while (std::scanf("%lf %lf %lf %lf %lf %lf %lf %lf", &lam, &mu, v, v + 1, v + 2, v + 3, v + 4, v + 5) == 8) {
    double sp[6];
    lsk::sigma_tension(v, lam, mu, sp);
    const double psi = lsk::strain_energy_tension(v, lam, mu);
    std::printf("%.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", sp[0], sp[1], sp[2], sp[3], sp[4], sp[5], psi);
  }
  return 0;
}
