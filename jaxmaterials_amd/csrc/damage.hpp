// Tension/compression split of the strain for the phase-field damage law
// (compute_sigma_damaged and compute_strain_energy,
// src/jaxmaterials/solver/lippmann_schwinger.py:36-100).
//
// The reference forms the symmetric matrix of a Voigt vector v as
//   M = [[v0, v5, v4], [v5, v1, v3], [v4, v3, v2]]
// (src/jaxmaterials/utilities.py:65-87) and maps matrices back the same way,
// so the split below works on that matrix (the xy and yz slots of the
// solver's Voigt order exchanged) and returns every component to its own
// slot, as the reference does.
//
// With eigenvalues l1 >= l2 >= l3 of M (FP64, trigonometric closed form;
// exact for diagonal M; near a repeated eigenvalue the closed form resolves
// the pair to ~1e-8 of |M|, far below the FP32 fields it feeds),
// the tension part M+ = sum_a max(l_a, 0) P_a follows from the spectral
// projectors without eigenvectors:
//   no positive eigenvalue: M+ = 0;  three: M+ = M;
//   one:  M+ = l1 (M - l2 I)(M - l3 I) / ((l1 - l2)(l1 - l3)),
//   two:  M+ = M - l3 (M - l1 I)(M - l2 I) / ((l3 - l1)(l3 - l2)),
// whose denominators are bounded away from zero by the sign change
// (l1 > 0 >= l2 in the first case, l2 > 0 >= l3 in the second) and whose
// prefactors l1 / (l1 - l2) and l3 / (l3 - l2) lie in [0, 1].
//   sigma+ = lambda max(tr, 0) I + 2 mu M+,
//   sigma  = g sigma+ + sigma-  =  C eps + (g - 1) sigma+,   g = (1 - d)^2 + k,
//   psi+   = lambda max(tr, 0)^2 / 2 + mu sum_a max(l_a, 0)^2.
#pragma once
#include <hip/hip_runtime.h>

#include <math.h>

namespace lsk {

struct Sym3 {
  double a00, a11, a22, a12, a02, a01;  // M from a Voigt vector (see above)
};

__device__ __forceinline__ Sym3 sym_of_voigt(const double* v) { return {v[0], v[1], v[2], v[3], v[4], v[5]}; }

// eigenvalues l1 >= l2 >= l3 of a symmetric 3 x 3 matrix
__device__ __forceinline__ void eig_sym3(const Sym3& m, double& l1, double& l2, double& l3) {
  const double p1 = m.a01 * m.a01 + m.a02 * m.a02 + m.a12 * m.a12;
  const double q = (m.a00 + m.a11 + m.a22) / 3.0;
  const double b00 = m.a00 - q, b11 = m.a11 - q, b22 = m.a22 - q;
  const double p2 = b00 * b00 + b11 * b11 + b22 * b22 + 2.0 * p1;
  if (p2 == 0.0) {
    l1 = l2 = l3 = q;
    return;
  }
  const double p = sqrt(p2 / 6.0);
  const double ip = 1.0 / p;
  const double c00 = b00 * ip, c11 = b11 * ip, c22 = b22 * ip;
  const double c12 = m.a12 * ip, c02 = m.a02 * ip, c01 = m.a01 * ip;
  const double det = c00 * (c11 * c22 - c12 * c12) - c01 * (c01 * c22 - c12 * c02) + c02 * (c01 * c12 - c11 * c02);
  double r = 0.5 * det;
  r = r < -1.0 ? -1.0 : (r > 1.0 ? 1.0 : r);
  const double phi = acos(r) / 3.0;
  l1 = q + 2.0 * p * cos(phi);
  l3 = q + 2.0 * p * cos(phi + 2.0943951023931954923);  // + 2 pi / 3
  l2 = 3.0 * q - l1 - l3;
  // keep the order exact against rounding
  if (l2 > l1) { const double t = l1; l1 = l2; l2 = t; }
  if (l3 > l2) { const double t = l2; l2 = l3; l3 = t; }
  if (l2 > l1) { const double t = l1; l1 = l2; l2 = t; }
}

// (M - a I)(M - b I) as a Voigt-slot vector (slots as sym_of_voigt)
__device__ __forceinline__ void shifted_product(const Sym3& m, double a, double b, double* o) {
  const double x00 = m.a00 - a, x11 = m.a11 - a, x22 = m.a22 - a;
  const double y00 = m.a00 - b, y11 = m.a11 - b, y22 = m.a22 - b;
  // rows of X = M - a I: (x00, a01, a02), (a01, x11, a12), (a02, a12, x22)
  o[0] = x00 * y00 + m.a01 * m.a01 + m.a02 * m.a02;
  o[1] = m.a01 * m.a01 + x11 * y11 + m.a12 * m.a12;
  o[2] = m.a02 * m.a02 + m.a12 * m.a12 + x22 * y22;
  o[3] = m.a01 * m.a02 + x11 * m.a12 + m.a12 * y22;  // (1, 2)
  o[4] = x00 * m.a02 + m.a01 * m.a12 + m.a02 * y22;  // (0, 2)
  o[5] = x00 * m.a01 + m.a01 * y11 + m.a02 * m.a12;  // (0, 1)
}

// M+ (Voigt slots) and the eigenvalues of M (unsorted when M is diagonal)
__device__ __forceinline__ void tension_part(const double* v, double* plus, double& l1, double& l2, double& l3) {
  const Sym3 m = sym_of_voigt(v);
  if (m.a12 == 0.0 && m.a02 == 0.0 && m.a01 == 0.0) {
    // diagonal (uniaxial and other axis-aligned states): exact
    plus[0] = fmax(m.a00, 0.0);
    plus[1] = fmax(m.a11, 0.0);
    plus[2] = fmax(m.a22, 0.0);
    plus[3] = plus[4] = plus[5] = 0.0;
    l1 = m.a00;
    l2 = m.a11;
    l3 = m.a22;
    return;
  }
  eig_sym3(m, l1, l2, l3);
  const int npos = (l1 > 0.0) + (l2 > 0.0) + (l3 > 0.0);
  if (npos == 0) {
#pragma unroll
    for (int c = 0; c < 6; ++c) plus[c] = 0.0;
  } else if (npos == 3) {
#pragma unroll
    for (int c = 0; c < 6; ++c) plus[c] = v[c];
  } else if (npos == 1) {
    double o[6];
    shifted_product(m, l2, l3, o);
    const double f = l1 / ((l1 - l2) * (l1 - l3));
#pragma unroll
    for (int c = 0; c < 6; ++c) plus[c] = f * o[c];
  } else {
    double o[6];
    shifted_product(m, l1, l2, o);
    const double f = l3 / ((l3 - l1) * (l3 - l2));
#pragma unroll
    for (int c = 0; c < 6; ++c) plus[c] = v[c] - f * o[c];
  }
}

// sigma+ = lambda max(tr, 0) I + 2 mu M+ (Voigt slots)
__device__ __forceinline__ void sigma_tension(const double* v, double lam, double mu, double* sp) {
  double plus[6], l1, l2, l3;
  tension_part(v, plus, l1, l2, l3);
  const double trp = fmax(v[0] + v[1] + v[2], 0.0);
#pragma unroll
  for (int c = 0; c < 6; ++c) sp[c] = 2.0 * mu * plus[c] + (c < 3 ? lam * trp : 0.0);
}

// psi+ = lambda max(tr, 0)^2 / 2 + mu sum max(l_a, 0)^2
__device__ __forceinline__ double strain_energy_tension(const double* v, double lam, double mu) {
  const Sym3 m = sym_of_voigt(v);
  double l1 = m.a00, l2 = m.a11, l3 = m.a22;
  if (m.a12 != 0.0 || m.a02 != 0.0 || m.a01 != 0.0) eig_sym3(m, l1, l2, l3);
  const double trp = fmax(v[0] + v[1] + v[2], 0.0);
  const double p1 = fmax(l1, 0.0), p2 = fmax(l2, 0.0), p3 = fmax(l3, 0.0);
  return 0.5 * lam * trp * trp + mu * (p1 * p1 + p2 * p2 + p3 * p3);
}

__device__ __forceinline__ double degradation(double d, double k) { return (1.0 - d) * (1.0 - d) + k; }

}  // namespace lsk
