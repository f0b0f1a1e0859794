// Reference fibre placement: Yang Chen's random fibre movement, with the
// reference's C ABI (src/jaxmaterials/data/libfibres.cc:31-179, bound by
// ctypes in src/jaxmaterials/data/distributions_fibres.py:121-170).
//
// Host code, built with g++ like the reference's own runtime compile
// (distributions_fibres.py:126-128).  The movement is a sequential Markov
// chain (each trial position depends on every fibre's current position and
// on one Mersenne-Twister stream), so it stays on the host; the GPU part of
// the reference placement is the rasteriser (microstructure.hip).
//
// Results are bit-identical to the reference's library for the same
// arguments and random state, which is what dataset parity needs:
//   * the random stream is std::mt19937 and its textual state
//     (libfibres.cc:61-74, 110-114, 170-176), the label permutation is
//     std::shuffle and the uniforms are std::uniform_real_distribution<double>
//     over [0, 1) (libfibres.cc:116-120), all from the same C++ standard
//     library, drawn in the same order (radius, then angle, per trial);
//   * the periodic distance is sqrt(min(4 L0 L1, d^2)) with d^2 the minimum
//     over the 3 x 3 periodic images (libfibres.cc:31-46).  Rounding is
//     monotone, so that minimum equals fl(min_x fl(dx^2) + min_y fl(dy^2)):
//     three candidates per axis instead of nine image pairs, same value;
//   * the three nearest neighbours come from a partial sort (the same three
//     values, summed in ascending order as after the full sort);
//   * the fibre's own entry carries the sentinel L0 + L1 (minus the radii
//     sum in the trial loop), as in libfibres.cc:131, 150.
// One difference: the reference loops forever if a sweep never ends without
// an unresolved overlap; here the sweeps stop after kMaxSweeps with a message.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <numeric>
#include <random>
#include <sstream>
#include <string>
#include <vector>

namespace {

constexpr unsigned kMaxSweeps = 1000000;

// Minimum over the offsets -L, 0, +L of fl(fl(p + o L) - q)^2.
inline double axis_min_sq(double p, double L, double q) {
  const double dm = (p - L) - q;
  const double d0 = p - q;
  const double dp = (p + L) - q;
  const double a = dm * dm, b = d0 * d0, c = dp * dp;
  double m = a < b ? a : b;
  return c < m ? c : m;
}

// Periodic distance between (px, py) and fibre j of the SoA coordinates.
inline double periodic_dist(double px, double py, double qx, double qy, double L0, double L1, double cap) {
  const double s = axis_min_sq(px, L0, qx) + axis_min_sq(py, L1, qy);
  return std::sqrt(s < cap ? s : cap);
}

void load_state(std::mt19937& rng, const char* state, unsigned length) {
  std::stringstream ss;
  ss.write(state, length);
  ss >> rng;
}

void store_state(const std::mt19937& rng, unsigned& length, char* state) {
  std::stringstream ss;
  ss << rng;
  const std::string s = ss.str();
  length = (unsigned)s.size();
  std::copy(s.begin(), s.end(), state);
}

// The fibre coordinates as two planes; written back to the interleaved
// caller array after every accepted move, like the reference.
struct Layout {
  std::vector<double> x, y;
};

// Clearances d_i = dist(p, q_i) - (r_p + r_i) of the fibres i in [lo, hi):
// adds the number of negative ones to `overlaps` and ORs into `below`
// whether any is below eps.  A branch-free loop that the compiler vectorises
// (the counts are kept in doubles, exact below 2^53, so that it does so with
// SSE2 too); an AVX2 clone is picked at load time where the CPU has it.
// Vector sqrt, min, add and multiply round exactly as the scalar ones.
__attribute__((target_clones("avx2", "default"))) void clearances(
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ r, size_t lo, size_t hi,
    double px, double py, double r_p, double L0, double L1, double cap, double eps, unsigned& overlaps,
    int& below) {
  double ov = 0.0, bl = 0.0;
  for (size_t i = lo; i < hi; ++i) {
    const double d = periodic_dist(px, py, x[i], y[i], L0, L1, cap) - (r_p + r[i]);
    ov += d < 0.0 ? 1.0 : 0.0;
    bl += d < eps ? 1.0 : 0.0;
  }
  overlaps += (unsigned)ov;
  below |= bl > 0.0;
}

}  // namespace

extern "C" {

// Periodic distances from p to the n points q (q[2j], q[2j+1]) in the box
// L[0] x L[1] (libfibres.cc:31-46).
void dist_periodic(const unsigned int n, const double p[2], const double* q, const double L[2], double* dist) {
  const double cap = 4.0 * L[0] * L[1];
  for (unsigned j = 0; j < n; ++j) dist[j] = periodic_dist(p[0], p[1], q[2 * j], q[2 * j + 1], L[0], L[1], cap);
}

// Textual state of std::mt19937 seeded with `seed` (libfibres.cc:61-74).
void initialise_rng(unsigned int seed, unsigned int* rng_state_length, char* rng_state) {
  std::mt19937 rng;
  rng.seed(seed);
  store_state(rng, *rng_state_length, rng_state);
}

// Random fibre movement (libfibres.cc:94-179): fibres are visited in a
// shuffled order; each is moved to a uniformly random point of the disc of
// radius (mean distance to its three nearest neighbours) around it until its
// clearance to every other fibre reaches eps_fibres, keeping the trial with
// fewest overlaps when it_max_ovlap trials fail.  Sweeps repeat num_repeats
// times and then until one ends with no failed fibre.  fibre_locations[2n]
// in/out; the random state in/out.
void move_fibres(const unsigned int n_fibres, const double* domain_size, const double* r_fibres,
                 double* fibre_locations, const unsigned int num_repeats, const double eps_fibres,
                 const unsigned int it_max_ovlap, unsigned int* rng_state_length, char* rng_state) {
  std::mt19937 rng;
  load_state(rng, rng_state, *rng_state_length);
  const unsigned n = n_fibres;
  std::vector<unsigned> order(n);
  std::iota(order.begin(), order.end(), 0u);
  std::shuffle(order.begin(), order.end(), rng);
  std::uniform_real_distribution<double> unit(0.0, 1.0);

  const double L0 = domain_size[0], L1 = domain_size[1];
  const double cap = 4.0 * L0 * L1;
  const double self_gap = L0 + L1;
  const double two_pi = 2 * M_PI;
  Layout f;
  f.x.resize(n);
  f.y.resize(n);
  for (unsigned i = 0; i < n; ++i) {
    f.x[i] = fibre_locations[2 * i];
    f.y[i] = fibre_locations[2 * i + 1];
  }
  // at least three entries: with fewer fibres the reference reads past its
  // array; here the missing neighbours count at the sentinel distance
  std::vector<double> gap(n < 3 ? 3 : n, self_gap);

  bool failed = true;
  unsigned sweep = 0;
  while (failed || sweep < num_repeats) {
    if (sweep >= kMaxSweeps) {
      fprintf(stderr, "move_fibres: overlaps unresolved after %u sweeps; stopping\n", sweep);
      break;
    }
    failed = false;
    for (unsigned v = 0; v < n; ++v) {
      const unsigned me = order[v];
      const double px = f.x[me], py = f.y[me];
      const double r_me = r_fibres[me];
      // trial radius: mean distance to the three nearest neighbours
      for (unsigned i = 0; i < n; ++i) gap[i] = periodic_dist(px, py, f.x[i], f.y[i], L0, L1, cap);
      gap[me] = self_gap;
      std::partial_sort(gap.begin(), gap.begin() + 3, gap.end());
      const double ur_sc = (gap[0] + gap[1] + gap[2]) / 3.0;

      // the reference loops while the smallest clearance is below eps_fibres,
      // i.e. while any clearance is (libfibres.cc:136, 153)
      bool crowded = 0.0 < eps_fibres;
      unsigned trials = 0;
      unsigned fewest = n + 1;
      double best_x = px, best_y = py;
      double nx = px, ny = py;
      while (crowded) {
        const double u_r = ur_sc * unit(rng);
        const double u_theta = two_pi * unit(rng);
        nx = px + u_r * std::cos(u_theta);
        ny = py + u_r * std::sin(u_theta);
        while (nx < 0) nx += L0;
        while (nx > L0) nx -= L0;
        while (ny < 0) ny += L1;
        while (ny > L1) ny -= L1;
        unsigned overlaps = 0;
        int below = 0;
        clearances(f.x.data(), f.y.data(), r_fibres, 0, me, nx, ny, r_me, L0, L1, cap, eps_fibres, overlaps, below);
        clearances(f.x.data(), f.y.data(), r_fibres, me + 1, n, nx, ny, r_me, L0, L1, cap, eps_fibres, overlaps,
                   below);
        const double own = self_gap - (r_me + r_me);
        overlaps += own < 0;
        below |= own < eps_fibres;
        crowded = below != 0;
        if (overlaps < fewest) {
          best_x = nx;
          best_y = ny;
          fewest = overlaps;
        }
        ++trials;
        if (trials > it_max_ovlap) {
          nx = best_x;
          ny = best_y;
          failed = true;
          break;
        }
      }
      f.x[me] = nx;
      f.y[me] = ny;
      fibre_locations[2 * me] = nx;
      fibre_locations[2 * me + 1] = ny;
    }
    ++sweep;
  }
  store_state(rng, *rng_state_length, rng_state);
}

}  // extern "C"
