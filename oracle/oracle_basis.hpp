// TEST INFRASTRUCTURE ONLY: CPU oracle for the TSEM operator path.
// Nothing in the product links or calls this file; tests/, the smoke check
// and bench.py's cpu_baseline leg use it as the checker.
//
// Polynomial layer: multi-index numbering, generalized Jacobi polynomials,
// collapsed Gauss-Legendre quadrature and the reference-tetrahedron basis,
// restated from TSEMLib:
//   include/Correspondence.h:64-109      (Correspondence<3>)
//   src/Miscellaneous.cpp:175-220, 510-542 (coefficient a, generalized Jacobi P)
//   src/QuadInfo.cpp:63-119              (build_quad_info)
//   src/BuildSpace.cpp:15-147, 295-433   (basis values, local transform)
// Basis gradients use forward-mode dual numbers through the same collapsed
// coordinate product instead of the reference's closed-form D coefficients
// (src/Miscellaneous.cpp:381-468); both give the exact derivative of the same
// polynomial, which the tests pin against finite differences.
#pragma once
#include <cmath>
#include <vector>
#include <array>
#include <cassert>

namespace oracle {

// ---- Correspondence<3> (include/Correspondence.h:64-109) -------------------
struct Corr {
  int M = 0, n = 0;
  int nm[3][64];                       // cumulative binomials
  std::vector<std::array<int, 3>> n2i; // number -> multi-index
  static int binom(int n_, int m) {
    int v = 1;
    if (n_ > 0 && n_ >= m) {
      if (n_ - m < m) m = n_ - m;
      for (int i = 0; i < m; ++i) { v *= (n_ - i); v /= (i + 1); }
    }
    return v;
  }
  void init(int M_) {
    M = M_;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j <= M; ++j) {
        nm[i][j] = binom(i + j, i);
        if (j > 0) nm[i][j] += nm[i][j - 1];
      }
    n = nm[2][M];
    n2i.assign(n, {-1, -1, -1});
    for (int a = 0; a <= M; ++a)
      for (int b = 0; a + b <= M; ++b)
        for (int c = 0; a + b + c <= M; ++c) {
          int k = i2n(a, b, c);
          assert(k >= 0 && k < n && n2i[k][0] < 0);
          n2i[k] = {a, b, c};
        }
  }
  int i2n(int a, int b, int c) const {
    int idx[3] = {a, b, c};
    int num = 0;
    for (int i = 0; i < 3; ++i) {
      int s = 0;
      for (int j = i; j < 3; ++j) s += idx[j];
      if (s == 0) break;
      num += nm[2 - i][s - 1];
    }
    return num;
  }
  int begin(int order) const { return order == 0 ? 0 : nm[2][order - 1]; }
  int end(int order) const { return nm[2][order]; }
};

// ---- dual numbers for gradients ---------------------------------------------
struct D3 {
  double v, d[3];
  D3(double x = 0) : v(x), d{0, 0, 0} {}
};
inline D3 operator+(const D3& a, const D3& b) { D3 r(a.v + b.v); for (int i = 0; i < 3; ++i) r.d[i] = a.d[i] + b.d[i]; return r; }
inline D3 operator-(const D3& a, const D3& b) { D3 r(a.v - b.v); for (int i = 0; i < 3; ++i) r.d[i] = a.d[i] - b.d[i]; return r; }
inline D3 operator*(const D3& a, const D3& b) { D3 r(a.v * b.v); for (int i = 0; i < 3; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i]; return r; }
inline D3 operator/(const D3& a, const D3& b) { D3 r(a.v / b.v); for (int i = 0; i < 3; ++i) r.d[i] = (a.d[i] * b.v - a.v * b.d[i]) / (b.v * b.v); return r; }
inline D3 operator*(double s, const D3& a) { D3 r(s * a.v); for (int i = 0; i < 3; ++i) r.d[i] = s * a.d[i]; return r; }
inline D3 operator-(const D3& a, double s) { D3 r(a); r.v -= s; return r; }
inline D3 operator-(double s, const D3& a) { D3 r(s - a.v); for (int i = 0; i < 3; ++i) r.d[i] = -a.d[i]; return r; }
inline D3 operator/(const D3& a, double s) { D3 r(a.v / s); for (int i = 0; i < 3; ++i) r.d[i] = a.d[i] / s; return r; }
inline D3 operator+(const D3& a, double s) { D3 r(a); r.v += s; return r; }
inline D3 operator+(double s, const D3& a) { D3 r(a); r.v += s; return r; }
inline double val(double x) { return x; }
inline double val(const D3& x) { return x.v; }

// ---- recurrence coefficient a (src/Miscellaneous.cpp:175-220) ---------------
inline double coef_a(int al, int be, int ind, int k) {
  if (k < 0) return 0.0;
  double ans = 0.0;
  switch (ind) {
    case 1:
      ans = (double)(2 * (k + 1) * (k + al + be + 1)) / ((2 * k + al + be + 1) * (2 * k + al + be + 2));
      if (al == -1 && be == -1) {
        if (k == 0) ans = 1; else if (k == 1) ans = 4; else if (k == 2) ans = 0.5;
        break;
      }
      if (k == 0 && al + be != -2) ans = 2.0 / (al + be + 2);
      break;
    case 2:
      if (0 <= k && k <= 2 && al == -1 && be == -1) { ans = 0; break; }
      if (k == 0 && al + be != -2) { ans = (double)(be - al) / (al + be + 2); break; }
      ans = (double)(be * be - al * al) / ((2 * k + al + be) * (2 * k + al + be + 2));
      break;
    case 3:
      ans = (double)(2 * (k + al) * (k + be)) / ((2 * k + al + be) * (2 * k + al + be + 1));
      if (al == -1 && be == -1) {
        if (k == 0) ans = 0; if (k == 1) ans = 1; if (k == 2) ans = 0;
        break;
      }
      if (k == 0 && al + be != -2) ans = 0;
      break;
  }
  return ans;
}

// generalized Jacobi polynomial, alpha,beta >= -1 (src/Miscellaneous.cpp:510-542)
template <class T>
T gjacobi(int al, int be, int k, T x) {
  if (k == 0) return T(1.0);
  if (al + be == -2) {
    if (k == 1) return x;
    return 0.25 * ((x - 1.0) * (x + 1.0)) * gjacobi<T>(1, 1, k - 2, x);
  }
  if (al == -1) return ((double)(k + be) / (2 * k)) * ((x - 1.0) * gjacobi<T>(1, be, k - 1, x));
  if (be == -1) return ((double)(k + al) / (2 * k)) * ((x + 1.0) * gjacobi<T>(al, 1, k - 1, x));
  T tp(1.0), ans(0.0);
  for (int j = 0; j <= k; ++j) {
    double f = 1.0;
    for (int i = 0; i < k - j; ++i) f *= (double)(al + j + 1 + i) / (i + 1);
    for (int i = 0; i < j; ++i) f *= (double)(k + al + be + 1 + i) / (i + 1);
    ans = ans + f * tp;
    tp = tp * ((x - 1.0) / 2.0);
  }
  return ans;
}

// J[0..n] of the (al,be) family at x by the three-term recurrence the basis
// construction uses (src/BuildSpace.cpp:29-45): J0=1, J1 = P_1, then a-coeffs.
template <class T>
void jac_rec(int al, int be, int n, T x, T* J) {
  J[0] = T(1.0);
  if (n >= 1) J[1] = gjacobi<T>(al, be, 1, x);
  for (int k = 1; k < n; ++k)
    J[k + 1] = ((x - coef_a(al, be, 2, k)) * J[k] - coef_a(al, be, 3, k) * J[k - 1]) / coef_a(al, be, 1, k);
}

// ---- quadrature (src/QuadInfo.cpp:63-119) -----------------------------------
// Gauss-Legendre on [-1,1] by Newton iteration on P_q.
inline void gauss_legendre(int q, std::vector<double>& x, std::vector<double>& w) {
  x.resize(q); w.resize(q);
  for (int i = 0; i < q; ++i) {
    double z = std::cos(M_PI * (i + 0.75) / (q + 0.5));
    for (int it = 0; it < 100; ++it) {
      double p0 = 1, p1 = z;
      for (int k = 2; k <= q; ++k) { double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k; p0 = p1; p1 = p2; }
      if (q == 1) { p1 = z; p0 = 1; }
      double dp = q * (z * p1 - p0) / (z * z - 1);
      double dz = p1 / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    double p0 = 1, p1 = z;
    for (int k = 2; k <= q; ++k) { double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k; p0 = p1; p1 = p2; }
    double dp = q * (z * p1 - p0) / (z * z - 1);
    x[q - 1 - i] = z;
    w[q - 1 - i] = 2.0 / ((1 - z * z) * dp * dp);
  }
}

struct Quad {
  int q = 0;
  // 1D: t in [0,1], bary (t, 1-t), weights sum 1
  std::vector<double> t1, w1;
  // 2D: bary (x, y, r=1-x-y), weights sum 1
  std::vector<std::array<double, 3>> p2;
  std::vector<double> w2;
  // 3D: reference tet points, weights sum 1
  std::vector<std::array<double, 3>> p3;
  std::vector<double> w3;
  void build(int q_) {
    q = q_;
    std::vector<double> x, w;
    gauss_legendre(q, x, w);
    t1.resize(q); w1.resize(q);
    for (int i = 0; i < q; ++i) { t1[i] = (1 + x[i]) * 0.5; w1[i] = w[i] * 0.5; }
    p2.resize(q * q); w2.resize(q * q);
    for (int i = 0; i < q; ++i)
      for (int j = 0; j < q; ++j) {
        int k = i * q + j;
        double a = t1[i] * (1 - t1[j]), b = t1[j];
        p2[k] = {a, b, 1 - a - b};
        w2[k] = 2 * (1 - t1[j]) * w1[i] * w1[j];
      }
    p3.resize((size_t)q * q * q); w3.resize((size_t)q * q * q);
    for (int i = 0; i < q; ++i)
      for (int j = 0; j < q; ++j)
        for (int l = 0; l < q; ++l) {
          size_t k = ((size_t)i * q + j) * q + l;
          p3[k] = {t1[i] * (1 - t1[j]) * (1 - t1[l]), t1[j] * (1 - t1[l]), t1[l]};
          w3[k] = 6 * (1 - t1[j]) * (1 - t1[l]) * (1 - t1[l]) * w1[i] * w1[j] * w1[l];
        }
  }
};

// ---- local transform J -> actual basis (src/BuildSpace.cpp:295-399) ---------
struct LocalTransform {
  std::vector<std::vector<std::pair<int, double>>> tl;  // [J index] -> (actual index, weight)
  void build(const Corr& c) {
    int M = c.M;
    tl.assign(c.n, {});
    tl[0] = {{0, 0.125}, {1, 0.125}, {2, 0.25}, {3, 0.5}};
    tl[1] = {{0, -0.5}, {1, 0.5}};
    tl[2] = {{0, -0.25}, {1, -0.25}, {2, 0.5}};
    tl[3] = {{0, -0.125}, {1, -0.125}, {2, -0.25}, {3, 0.5}};
    for (int l = 2; l <= M; ++l) {
      int i3 = c.i2n(l, 0, 0);
      tl[i3] = {{i3, 2.0}};
      int i1 = c.i2n(0, l, 0), i2 = c.i2n(1, l - 1, 0);
      tl[i1] = {{i1, 1.0}, {i2, 1.0}};
      tl[i2] = {{i1, (double)(l - 1) / l}, {i2, (double)(-(l - 1)) / l}};
      i1 = c.i2n(0, 0, l); i2 = c.i2n(1, 0, l - 1); i3 = c.i2n(0, 1, l - 1);
      tl[i1] = {{i1, 0.5}, {i2, 0.5}, {i3, 1.0}};
      tl[i2] = {{i1, (double)(l - 1) / l}, {i2, (double)(-(l - 1)) / l}};
      tl[i3] = {{i1, (double)(l - 1) / (2 * l)}, {i2, (double)(l - 1) / (2 * l)}, {i3, (double)(-(l - 1)) / l}};
    }
    for (int l1 = 2; l1 <= M; ++l1)
      for (int l2 = 1; l2 <= M - l1; ++l2) {
        int i1 = c.i2n(0, l1, l2), i2 = c.i2n(1, l1 - 1, l2);
        tl[i1] = {{i1, 1.0}, {i2, 1.0}};
        tl[i2] = {{i1, (double)(l1 - 1) / l1}, {i2, (double)(-(l1 - 1)) / l1}};
        int a = c.i2n(l1, 0, l2), b = c.i2n(l1, l2, 0);
        tl[a] = {{a, 2.0}};
        tl[b] = {{b, 2.0}};
      }
    for (int l1 = 2; l1 <= M; ++l1)
      for (int l2 = 1; l2 <= M - l1; ++l2)
        for (int l3 = 1; l3 <= M - l1 - l2; ++l3) {
          int k = c.i2n(l1, l2, l3);
          tl[k] = {{k, 1.0}};
        }
  }
};

// Jacobi-product basis J_alpha at a reference-tet point (src/BuildSpace.cpp:22-52)
template <class T>
void jproduct(const Corr& c, T x, T y, T z, std::vector<T>& out) {
  int M = c.M;
  out.assign(c.n, T(0.0));
  T t1 = 1.0 - (y + z), t2 = 1.0 - z;
  T xi = 2.0 * x / t1 - 1.0, eta = 2.0 * y / t2 - 1.0, zeta = 2.0 * z - 1.0;
  std::vector<T> Jxi(M + 2), Jeta(M + 2), Jzeta(M + 2);
  jac_rec<T>(-1, -1, M, xi, Jxi.data());
  std::vector<T> pt1(M + 1), pt2(M + 1);
  pt1[0] = T(1.0); pt2[0] = T(1.0);
  for (int l = 1; l <= M; ++l) { pt1[l] = pt1[l - 1] * t1; pt2[l] = pt2[l - 1] * t2; }
  for (int l1 = 0; l1 <= M; ++l1) {
    int a2 = 2 * l1 - 1;
    jac_rec<T>(a2, -1, M - l1, eta, Jeta.data());
    for (int l2 = 0; l2 <= M - l1; ++l2) {
      int a3 = 2 * l1 + 2 * l2 - 1;
      jac_rec<T>(a3, -1, M - l1 - l2, zeta, Jzeta.data());
      for (int l3 = 0; l3 <= M - l1 - l2; ++l3)
        out[c.i2n(l1, l2, l3)] = pt1[l1] * Jxi[l1] * (pt2[l2] * Jeta[l2]) * Jzeta[l3];
    }
  }
}

// actual basis (vertex/edge/face/interior modes) at a point
template <class T>
void actual_basis(const Corr& c, const LocalTransform& lt, T x, T y, T z, std::vector<T>& out) {
  std::vector<T> J;
  jproduct<T>(c, x, y, z, J);
  out.assign(c.n, T(0.0));
  for (int i = 0; i < c.n; ++i)
    for (auto& pr : lt.tl[i]) out[pr.first] = out[pr.first] + pr.second * J[i];
}

}  // namespace oracle
