// Reference tetrahedron of the hierarchical Jacobi basis: numbering of modes,
// basis evaluation, quadrature and the reference matrices the device kernels
// use.  Host code, long double internally.
//
// Basis (TSEMLib src/BuildSpace.cpp:15-52, 295-399): for a multi-index
// (l1,l2,l3) the Jacobi product
//   J = (1-y-z)^l1 P_l1^{-1,-1}(xi) (1-z)^l2 P_l2^{2l1-1,-1}(eta) P_l3^{2l1+2l2-1,-1}(zeta)
// with collapsed coordinates xi = 2x/(1-y-z)-1, eta = 2y/(1-z)-1, zeta = 2z-1,
// and the "actual" vertex/edge/face/interior modes are fixed sparse linear
// combinations of these products.  Here P^{a,b} with a or b = -1 follows the
// generalized definition of src/Miscellaneous.cpp:510-542; the ordinary
// families are evaluated with the textbook three-term recurrence.
#pragma once
#include <array>
#include <vector>

namespace tsem {

// multi-index numbering in degree blocks (include/Correspondence.h:64-109):
// all |a| = 0, then |a| = 1 in order (100),(010),(001), then |a| = 2 ...
struct ModeIndex {
  int M = 0, K = 0;
  std::vector<std::array<int, 3>> mi;  // number -> multi-index
  void init(int M_);
  int number(int a, int b, int c) const;  // multi-index -> number
  static int count_upto(int d) { return d < 0 ? 0 : (d + 1) * (d + 2) * (d + 3) / 6; }
};

// Gauss-Legendre rule on [0,1], weights summing to 1
void gauss_legendre01(int q, std::vector<long double>& t, std::vector<long double>& w);

struct RefElement {
  int M = 0, K = 0, Kd = 0;  // Kd = modes of degree <= M-1
  ModeIndex idx;
  // actual = sum over Jacobi products: for each actual mode, list of (J index, weight)
  std::vector<std::vector<std::pair<int, long double>>> comb;
  // reference matrices (row-major K x K), unit tetrahedron
  std::vector<double> mass;        // int phi_i phi_j
  std::vector<double> stiff[6];    // int (D_k phi_i)(D_k phi_j), D_k in the reference's pair order
  std::vector<double> grad2[9];    // int d_a phi_i d_b phi_j
  // derivative factors: dfac[a][m][i], m < Kd: coefficients of d_a phi_i in an
  // orthonormal basis of P_{M-1}; dfac[a]^T dfac[b] = grad2[3a+b]
  std::vector<double> dfac[3];
  // mass factor: mfac[m][i], mfac^T mfac = mass
  std::vector<double> mfac;
  void build(int M_);
  // the reference element of order M, built once per process (thread safe)
  static const RefElement& cached(int M);
  // actual basis values (and gradients) at a reference point
  void eval(long double x, long double y, long double z, long double* val, long double* grad /*K*3 or null*/) const;
};

}  // namespace tsem
