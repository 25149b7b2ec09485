// Reference element construction (see refel.hpp).
#include "refel.hpp"

#include <map>
#include <memory>
#include <mutex>

#include <cmath>
#include <stdexcept>

namespace tsem {

using LD = long double;

void ModeIndex::init(int M_) {
  M = M_;
  K = count_upto(M);
  mi.assign(K, {0, 0, 0});
  int n = 0;
  for (int d = 0; d <= M; ++d)          // degree blocks
    for (int a = d; a >= 0; --a)        // first component descending
      for (int b = d - a; b >= 0; --b)  // second descending, third fills
        mi[n++] = {a, b, d - a - b};
}

int ModeIndex::number(int a, int b, int c) const {
  // block offset + position inside the block: within degree d the order is
  // a descending, then b descending
  int d = a + b + c;
  int off = count_upto(d - 1);
  int pos = 0;
  for (int aa = d; aa > a; --aa) pos += d - aa + 1;
  pos += (d - a) - b;
  return off + pos;
}

void gauss_legendre01(int q, std::vector<LD>& t, std::vector<LD>& w) {
  t.resize(q); w.resize(q);
  const LD pi = 3.141592653589793238462643383279502884L;
  for (int i = 0; i < q; ++i) {
    LD z = cosl(pi * (i + 0.75L) / (q + 0.5L));
    LD dp = 1;
    for (int it = 0; it < 60; ++it) {
      LD p0 = 1, p1 = z;
      for (int k = 2; k <= q; ++k) { LD p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k; p0 = p1; p1 = p2; }
      if (q == 1) { p1 = z; p0 = 1; }
      dp = q * (z * p1 - p0) / (z * z - 1);
      LD dz = p1 / dp;
      z -= dz;
      if (fabsl(dz) < 1e-19L) break;
    }
    LD p0 = 1, p1 = z;
    for (int k = 2; k <= q; ++k) { LD p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k; p0 = p1; p1 = p2; }
    dp = q * (z * p1 - p0) / (z * z - 1);
    t[q - 1 - i] = (1 + z) / 2;
    w[q - 1 - i] = 1 / ((1 - z * z) * dp * dp);
  }
}

namespace {

// value + gradient in (x,y,z)
struct Dl {
  LD v, g[3];
  Dl(LD x = 0) : v(x), g{0, 0, 0} {}
};
Dl operator+(const Dl& a, const Dl& b) { Dl r(a.v + b.v); for (int i = 0; i < 3; ++i) r.g[i] = a.g[i] + b.g[i]; return r; }
Dl operator-(const Dl& a, const Dl& b) { Dl r(a.v - b.v); for (int i = 0; i < 3; ++i) r.g[i] = a.g[i] - b.g[i]; return r; }
Dl operator*(const Dl& a, const Dl& b) { Dl r(a.v * b.v); for (int i = 0; i < 3; ++i) r.g[i] = a.g[i] * b.v + a.v * b.g[i]; return r; }
Dl operator*(LD s, const Dl& a) { Dl r(s * a.v); for (int i = 0; i < 3; ++i) r.g[i] = s * a.g[i]; return r; }
Dl operator/(const Dl& a, const Dl& b) {
  Dl r(a.v / b.v);
  for (int i = 0; i < 3; ++i) r.g[i] = (a.g[i] * b.v - a.v * b.g[i]) / (b.v * b.v);
  return r;
}

// P_n^{(a,b)} for a,b > -1, textbook recurrence, n = 0..N into out
void jacobi_std(LD a, LD b, int N, const Dl& x, Dl* out) {
  out[0] = Dl(1);
  if (N >= 1) out[1] = Dl((a - b) / 2) + ((a + b + 2) / 2) * x;
  for (int n = 2; n <= N; ++n) {
    LD c = 2 * n + a + b;
    LD A1 = 2 * n * (n + a + b) * (c - 2);
    LD A2 = (c - 1) * (a * a - b * b);
    LD A3 = (c - 1) * c * (c - 2);
    LD A4 = 2 * (n + a - 1) * (n + b - 1) * c;
    out[n] = (1 / A1) * ((Dl(A2) + A3 * x) * out[n - 1] - A4 * out[n - 2]);
  }
}

// generalized family with a or b = -1 allowed (src/Miscellaneous.cpp:510-542)
void jacobi_gen(int a, int b, int N, const Dl& x, Dl* out) {
  if (a > -1 && b > -1) { jacobi_std(a, b, N, x, out); return; }
  std::vector<Dl> tmp(N + 2);
  out[0] = Dl(1);
  if (a == -1 && b == -1) {
    if (N >= 1) out[1] = x;
    if (N >= 2) jacobi_std(1, 1, N - 2, x, tmp.data());
    for (int n = 2; n <= N; ++n) out[n] = 0.25L * ((x - Dl(1)) * (x + Dl(1))) * tmp[n - 2];
    return;
  }
  if (b == -1) {  // a > -1
    if (N >= 1) jacobi_std(a, 1, N - 1, x, tmp.data());
    for (int n = 1; n <= N; ++n) out[n] = ((LD)(n + a) / (2 * n)) * ((x + Dl(1)) * tmp[n - 1]);
    return;
  }
  // a == -1, b > -1
  if (N >= 1) jacobi_std(1, b, N - 1, x, tmp.data());
  for (int n = 1; n <= N; ++n) out[n] = ((LD)(n + b) / (2 * n)) * ((x - Dl(1)) * tmp[n - 1]);
}

// all Jacobi products at a point, indexed by mode number
void products(const ModeIndex& idx, const Dl& x, const Dl& y, const Dl& z, std::vector<Dl>& J) {
  int M = idx.M;
  J.assign(idx.K, Dl(0));
  Dl one(1);
  Dl B = one - y - z, C = one - z;
  Dl xi = (2.0L * x) / B - one, eta = (2.0L * y) / C - one, zeta = 2.0L * z - one;
  std::vector<Dl> Pxi(M + 1), Peta(M + 1), Pzeta(M + 1), Bp(M + 1), Cp(M + 1);
  jacobi_gen(-1, -1, M, xi, Pxi.data());
  Bp[0] = one; Cp[0] = one;
  for (int l = 1; l <= M; ++l) { Bp[l] = Bp[l - 1] * B; Cp[l] = Cp[l - 1] * C; }
  for (int l1 = 0; l1 <= M; ++l1) {
    jacobi_gen(2 * l1 - 1, -1, M - l1, eta, Peta.data());
    Dl f1 = Bp[l1] * Pxi[l1];
    for (int l2 = 0; l1 + l2 <= M; ++l2) {
      jacobi_gen(2 * l1 + 2 * l2 - 1, -1, M - l1 - l2, zeta, Pzeta.data());
      Dl f12 = f1 * (Cp[l2] * Peta[l2]);
      for (int l3 = 0; l1 + l2 + l3 <= M; ++l3) J[idx.number(l1, l2, l3)] = f12 * Pzeta[l3];
    }
  }
}

void cholesky(std::vector<LD>& A, int n) {  // in place, lower
  for (int j = 0; j < n; ++j) {
    LD s = A[(size_t)j * n + j];
    for (int k = 0; k < j; ++k) s -= A[(size_t)j * n + k] * A[(size_t)j * n + k];
    if (!(s > 0)) throw std::runtime_error("reference mass matrix not positive definite");
    LD d = sqrtl(s);
    A[(size_t)j * n + j] = d;
    for (int i = j + 1; i < n; ++i) {
      LD t = A[(size_t)i * n + j];
      for (int k = 0; k < j; ++k) t -= A[(size_t)i * n + k] * A[(size_t)j * n + k];
      A[(size_t)i * n + j] = t / d;
    }
    for (int i = 0; i < j; ++i) A[(size_t)i * n + j] = 0;
  }
}

}  // namespace

void RefElement::eval(LD x, LD y, LD z, LD* val, LD* grad) const {
  Dl X(x), Y(y), Z(z);
  X.g[0] = 1; Y.g[1] = 1; Z.g[2] = 1;
  std::vector<Dl> J;
  products(idx, X, Y, Z, J);
  for (int k = 0; k < K; ++k) {
    Dl s(0);
    for (auto& pr : comb[k]) s = s + pr.second * J[pr.first];
    if (val) val[k] = s.v;
    if (grad) for (int d = 0; d < 3; ++d) grad[3 * k + d] = s.g[d];
  }
}

void RefElement::build(int M_) {
  if (M_ < 1 || M_ > 12) throw std::runtime_error("polynomial order out of range 1..12");
  M = M_;
  idx.init(M);
  K = idx.K;
  Kd = ModeIndex::count_upto(M - 1);
  // actual modes as combinations of Jacobi products
  comb.assign(K, {});
  auto n = [&](int a, int b, int c) { return idx.number(a, b, c); };
  // vertex modes: the barycentric coordinates
  comb[n(0, 0, 0)] = {{n(0, 0, 0), 0.125L}, {n(1, 0, 0), -0.5L}, {n(0, 1, 0), -0.25L}, {n(0, 0, 1), -0.125L}};
  comb[n(1, 0, 0)] = {{n(0, 0, 0), 0.125L}, {n(1, 0, 0), 0.5L}, {n(0, 1, 0), -0.25L}, {n(0, 0, 1), -0.125L}};
  comb[n(0, 1, 0)] = {{n(0, 0, 0), 0.25L}, {n(0, 1, 0), 0.5L}, {n(0, 0, 1), -0.25L}};
  comb[n(0, 0, 1)] = {{n(0, 0, 0), 0.5L}, {n(0, 0, 1), 0.5L}};
  for (int l = 2; l <= M; ++l) {
    LD r = (LD)(l - 1) / l;
    comb[n(l, 0, 0)] = {{n(l, 0, 0), 2.0L}};
    comb[n(0, l, 0)] = {{n(0, l, 0), 1.0L}, {n(1, l - 1, 0), r}};
    comb[n(1, l - 1, 0)] = {{n(0, l, 0), 1.0L}, {n(1, l - 1, 0), -r}};
    comb[n(0, 0, l)] = {{n(0, 0, l), 0.5L}, {n(1, 0, l - 1), r}, {n(0, 1, l - 1), r / 2}};
    comb[n(1, 0, l - 1)] = {{n(0, 0, l), 0.5L}, {n(1, 0, l - 1), -r}, {n(0, 1, l - 1), r / 2}};
    comb[n(0, 1, l - 1)] = {{n(0, 0, l), 1.0L}, {n(0, 1, l - 1), -r}};
  }
  for (int l1 = 2; l1 <= M; ++l1)
    for (int l2 = 1; l1 + l2 <= M; ++l2) {
      LD r = (LD)(l1 - 1) / l1;
      comb[n(0, l1, l2)] = {{n(0, l1, l2), 1.0L}, {n(1, l1 - 1, l2), r}};
      comb[n(1, l1 - 1, l2)] = {{n(0, l1, l2), 1.0L}, {n(1, l1 - 1, l2), -r}};
      comb[n(l1, 0, l2)] = {{n(l1, 0, l2), 2.0L}};
      comb[n(l1, l2, 0)] = {{n(l1, l2, 0), 2.0L}};
    }
  for (int l1 = 2; l1 <= M; ++l1)
    for (int l2 = 1; l1 + l2 <= M; ++l2)
      for (int l3 = 1; l1 + l2 + l3 <= M; ++l3) comb[n(l1, l2, l3)] = {{n(l1, l2, l3), 1.0L}};
  for (int k = 0; k < K; ++k)
    if (comb[k].empty()) throw std::runtime_error("mode without definition");

  // exact collapsed Gauss-Legendre quadrature of degree >= 2M+2 per direction
  int q = M + 3;
  std::vector<LD> t, w;
  gauss_legendre01(q, t, w);
  std::vector<LD> Mm((size_t)K * K, 0), G2[9];
  for (int k = 0; k < 9; ++k) G2[k].assign((size_t)K * K, 0);
  std::vector<LD> val(K), grd(3 * K);
  for (int i = 0; i < q; ++i)
    for (int j = 0; j < q; ++j)
      for (int l = 0; l < q; ++l) {
        LD x = t[i] * (1 - t[j]) * (1 - t[l]), y = t[j] * (1 - t[l]), z = t[l];
        LD wt = w[i] * w[j] * w[l] * (1 - t[j]) * (1 - t[l]) * (1 - t[l]);  // unit-tet measure
        eval(x, y, z, val.data(), grd.data());
        for (int a = 0; a < K; ++a) {
          LD va = wt * val[a];
          for (int b = 0; b < K; ++b) Mm[(size_t)a * K + b] += va * val[b];
          for (int da = 0; da < 3; ++da) {
            LD ga = wt * grd[3 * a + da];
            for (int db = 0; db < 3; ++db) {
              LD* row = &G2[3 * da + db][(size_t)a * K];
              for (int b = 0; b < K; ++b) row[b] += ga * grd[3 * b + db];
            }
          }
        }
      }
  mass.assign((size_t)K * K, 0);
  for (size_t s = 0; s < (size_t)K * K; ++s) mass[s] = (double)Mm[s];
  for (int k = 0; k < 9; ++k) {
    grad2[k].assign((size_t)K * K, 0);
    for (size_t s = 0; s < (size_t)K * K; ++s) grad2[k][s] = (double)G2[k][s];
  }
  // pair matrices in the reference's order: d1, d2-d1, d1-d3, d2, d3-d2, d3
  auto G = [&](int a, int b, size_t s) { return G2[3 * a + b][s]; };
  for (int k = 0; k < 6; ++k) stiff[k].assign((size_t)K * K, 0);
  for (size_t s = 0; s < (size_t)K * K; ++s) {
    stiff[0][s] = (double)G(0, 0, s);
    stiff[1][s] = (double)(G(1, 1, s) - G(1, 0, s) - G(0, 1, s) + G(0, 0, s));
    stiff[2][s] = (double)(G(0, 0, s) - G(0, 2, s) - G(2, 0, s) + G(2, 2, s));
    stiff[3][s] = (double)G(1, 1, s);
    stiff[4][s] = (double)(G(2, 2, s) - G(2, 1, s) - G(1, 2, s) + G(1, 1, s));
    stiff[5][s] = (double)G(2, 2, s);
  }
  // orthonormal factorizations: mass = L L^T (K x K); leading Kd block gives
  // the orthonormal basis of P_{M-1}; d_a phi_i = sum_m dfac[a][m][i] psi_m
  std::vector<LD> L(Mm);
  cholesky(L, K);
  mfac.assign((size_t)K * K, 0);
  for (int m = 0; m < K; ++m)
    for (int i = 0; i < K; ++i) mfac[(size_t)m * K + i] = (double)L[(size_t)i * K + m];
  // B_a[m][i] = int phi_m d_a phi_i for m < Kd
  std::vector<LD> Bq[3];
  for (int a = 0; a < 3; ++a) Bq[a].assign((size_t)Kd * K, 0);
  for (int i = 0; i < q; ++i)
    for (int j = 0; j < q; ++j)
      for (int l = 0; l < q; ++l) {
        LD x = t[i] * (1 - t[j]) * (1 - t[l]), y = t[j] * (1 - t[l]), z = t[l];
        LD wt = w[i] * w[j] * w[l] * (1 - t[j]) * (1 - t[l]) * (1 - t[l]);
        eval(x, y, z, val.data(), grd.data());
        for (int m = 0; m < Kd; ++m)
          for (int a = 0; a < 3; ++a)
            for (int ii = 0; ii < K; ++ii) Bq[a][(size_t)m * K + ii] += wt * val[m] * grd[3 * ii + a];
      }
  // dfac = L_d^{-1} B (forward substitution with the leading Kd block of L)
  for (int a = 0; a < 3; ++a) {
    std::vector<LD> D(Bq[a]);
    for (int ii = 0; ii < K; ++ii)
      for (int m = 0; m < Kd; ++m) {
        LD s = D[(size_t)m * K + ii];
        for (int k = 0; k < m; ++k) s -= L[(size_t)m * K + k] * D[(size_t)k * K + ii];
        D[(size_t)m * K + ii] = s / L[(size_t)m * K + m];
      }
    dfac[a].assign((size_t)Kd * K, 0);
    for (size_t s = 0; s < (size_t)Kd * K; ++s) dfac[a][s] = (double)D[s];
  }
}

const RefElement& RefElement::cached(int M) {
  static std::mutex mu;
  static std::map<int, std::unique_ptr<RefElement>> cache;
  std::lock_guard<std::mutex> lock(mu);
  auto& slot = cache[M];
  if (!slot) {
    slot.reset(new RefElement());
    slot->build(M);
  }
  return *slot;
}

}  // namespace tsem
