// Eigen solvers on the device operators of a space: block LOBPCG and inverse
// power iteration for A x = lambda B x with A = stiff S + mass M (+ potential
// M_V, the mass matrix weighted by the potential set on the space) and B = M,
// both with the space's Dirichlet elimination (vectors kept zero on the
// eliminated DOFs, so the eliminated rows add no spurious eigenpairs).
//
// LOBPCG follows the block method of TSEMSolver::solve_LOBPCG
// (src/TSEMSolver.cpp:2224-2407): trial block [X W P], Gram matrices of A
// and B, Rayleigh-Ritz, X and P recombined from the trial block, stop on the
// change of the eigenvalues.  Corrected here: the reference solves a fixed
// 3 x 3 problem for one pair (dsygvx with N = 3, il = iu = 1, :2349) and
// leaves the other eigenvalues unset; this solves the full 3m x 3m problem
// for all m pairs.  W is the preconditioned residual (Jacobi, or one FP32
// p-multigrid V-cycle per column) instead of an inner PCG solve of the
// rebuilt matrix S/2 - lambda M (:2292-2323).  For stability near
// convergence, converged pairs leave W (soft locking), W is B-orthogonalized
// against X and P, and W and P are B-orthonormalized with dependent
// directions dropped.  Gram matrices, the Rayleigh-Ritz problem and the
// recombinations are FP64.
//
// Inverse power follows solve_InversePower (src/TSEMSolver.cpp:2041-2069):
// x <- A^-1 B x by PCG, B-normalize, e = x^T A x, stop when |e - e_last| < tol.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <numeric>
#include <vector>

#include "handle.hpp"

using namespace tsem;

namespace {

// Symmetric eigenproblem (n x n, row-major): eigenvalues ascending in w,
// eigenvectors in the columns of V (row-major).  Householder reduction to
// tridiagonal form with the transforms accumulated, then the implicit QL
// method with Wilkinson shifts on the tridiagonal matrix (Golub and Van Loan,
// Matrix Computations, sections 8.3.1 and 8.3.3).  The accumulated transform
// is kept transposed (row j = basis vector j) so every rotation and every
// Householder update runs along contiguous rows.
void sym_eig(int n, std::vector<double> A, std::vector<double>& w, std::vector<double>& V) {
  std::vector<double> Zt((size_t)n * n, 0.0), d(n), e(n, 0.0), v(n), p(n), u(n);
  for (int i = 0; i < n; ++i) Zt[(size_t)i * n + i] = 1.0;
  for (int k = 0; k + 2 < n; ++k) {
    // H = I - beta v v^T maps A[k+1:, k] onto a multiple of e_{k+1}
    double s = 0.0;
    for (int i = k + 1; i < n; ++i) s += A[(size_t)i * n + k] * A[(size_t)i * n + k];
    const double x0 = A[(size_t)(k + 1) * n + k];
    const double tail = s - x0 * x0;
    if (tail == 0.0) continue;  // column already tridiagonal
    const double alpha = x0 > 0 ? -std::sqrt(s) : std::sqrt(s);
    for (int i = k + 1; i < n; ++i) v[i] = A[(size_t)i * n + k];
    v[k + 1] = x0 - alpha;
    const double beta = 2.0 / (tail + v[k + 1] * v[k + 1]);
    // trailing block <- H A H = A - v q^T - q v^T with p = beta A v and
    // q = p - (beta p.v / 2) v
    double pv = 0.0;
    for (int i = k + 1; i < n; ++i) {
      const double* Ai = &A[(size_t)i * n];
      double acc = 0.0;
      for (int j = k + 1; j < n; ++j) acc += Ai[j] * v[j];
      p[i] = beta * acc;
      pv += p[i] * v[i];
    }
    const double h = 0.5 * beta * pv;
    for (int i = k + 1; i < n; ++i) p[i] -= h * v[i];
    for (int i = k + 1; i < n; ++i) {
      double* Ai = &A[(size_t)i * n];
      for (int j = k + 1; j < n; ++j) Ai[j] -= v[i] * p[j] + p[i] * v[j];
    }
    A[(size_t)(k + 1) * n + k] = A[(size_t)k * n + k + 1] = alpha;
    for (int i = k + 2; i < n; ++i) A[(size_t)i * n + k] = A[(size_t)k * n + i] = 0.0;
    // accumulated transform: Zt[k+1:, :] <- H Zt[k+1:, :]
    std::fill(u.begin(), u.end(), 0.0);
    for (int i = k + 1; i < n; ++i) {
      const double* Zi = &Zt[(size_t)i * n];
      for (int c = 0; c < n; ++c) u[c] += v[i] * Zi[c];
    }
    for (int i = k + 1; i < n; ++i) {
      const double f = beta * v[i];
      double* Zi = &Zt[(size_t)i * n];
      for (int c = 0; c < n; ++c) Zi[c] -= f * u[c];
    }
  }
  // tridiagonal: d on the diagonal, e[i] couples rows i and i + 1
  for (int i = 0; i < n; ++i) d[i] = A[(size_t)i * n + i];
  for (int i = 0; i + 1 < n; ++i) e[i] = A[(size_t)(i + 1) * n + i];
  const double eps = 2.220446049250313e-16;
  for (int l = 0; l < n; ++l) {
    for (int iter = 0; iter < 64; ++iter) {
      // the first negligible coupling at or after l closes the active block
      int m = l;
      for (; m + 1 < n; ++m)
        if (std::fabs(e[m]) <= eps * (std::fabs(d[m]) + std::fabs(d[m + 1]))) break;
      if (m == l) break;
      // shift: the eigenvalue of the leading 2 x 2 block nearer d[l]
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = std::hypot(g, 1.0);
      g = d[m] - d[l] + e[l] / (g + (g >= 0 ? r : -r));
      double s = 1.0, c = 1.0, shift = 0.0;
      bool deflated = false;
      for (int i = m - 1; i >= l; --i) {
        // Givens rotation of rows i, i + 1 chasing the bulge upwards
        const double f = s * e[i], b = c * e[i];
        r = std::hypot(f, g);
        e[i + 1] = r;
        if (r == 0.0) {  // the coupling underflowed: split the block here
          d[i + 1] -= shift;
          e[m] = 0.0;
          deflated = true;
          break;
        }
        s = f / r;
        c = g / r;
        g = d[i + 1] - shift;
        r = (d[i] - g) * s + 2.0 * c * b;
        shift = s * r;
        d[i + 1] = g + shift;
        g = c * r - b;
        double* Za = &Zt[(size_t)i * n];
        double* Zb = &Zt[(size_t)(i + 1) * n];
        for (int k = 0; k < n; ++k) {
          const double za = Za[k], zb = Zb[k];
          Zb[k] = s * za + c * zb;
          Za[k] = c * za - s * zb;
        }
      }
      if (deflated) continue;
      d[l] -= shift;
      e[l] = g;
      e[m] = 0.0;
    }
  }
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return d[a] < d[b]; });
  w.resize(n);
  V.assign((size_t)n * n, 0.0);
  for (int c = 0; c < n; ++c) {
    w[c] = d[order[c]];
    const double* Zc = &Zt[(size_t)order[c] * n];
    for (int k = 0; k < n; ++k) V[(size_t)k * n + c] = Zc[k];
  }
}

// Rayleigh-Ritz on the trial space: the m lowest eigenpairs of
// GA c = lambda GB c (p x p, symmetric; GB positive semidefinite).  GB is
// equilibrated by its diagonal and directions whose relative GB-eigenvalue
// is below `drop` are removed, so a nearly dependent trial block is handled
// as a pseudo-inverse.  Returns C (p x m, row-major, GB-orthonormal columns)
// and the eigenvalues; false when fewer than m directions remain.
bool rayleigh_ritz(int p, std::vector<double> GA, std::vector<double> GB, int m, double drop, std::vector<double>& C,
                   std::vector<double>& lam) {
  for (int i = 0; i < p; ++i)
    for (int j = i + 1; j < p; ++j) {
      double a = 0.5 * (GA[(size_t)i * p + j] + GA[(size_t)j * p + i]);
      double b = 0.5 * (GB[(size_t)i * p + j] + GB[(size_t)j * p + i]);
      GA[(size_t)i * p + j] = GA[(size_t)j * p + i] = a;
      GB[(size_t)i * p + j] = GB[(size_t)j * p + i] = b;
    }
  std::vector<double> d(p);
  for (int i = 0; i < p; ++i) {
    double g = GB[(size_t)i * p + i];
    d[i] = g > 0 ? 1.0 / std::sqrt(g) : 0.0;
  }
  for (int i = 0; i < p; ++i)
    for (int j = 0; j < p; ++j) {
      GB[(size_t)i * p + j] *= d[i] * d[j];
      GA[(size_t)i * p + j] *= d[i] * d[j];
    }
  std::vector<double> wb, Vb;
  sym_eig(p, GB, wb, Vb);
  double wmax = wb.empty() ? 0.0 : wb.back();
  std::vector<int> keep;
  for (int k = 0; k < p; ++k)
    if (wb[k] > drop * wmax) keep.push_back(k);
  int r = (int)keep.size();
  if (r < m) return false;
  // Z = Vb[:, keep] / sqrt(wb): GB-orthonormal basis of the kept directions
  std::vector<double> Z((size_t)p * r);
  for (int i = 0; i < p; ++i)
    for (int c = 0; c < r; ++c) Z[(size_t)i * r + c] = Vb[(size_t)i * p + keep[c]] / std::sqrt(wb[keep[c]]);
  // H = Z^T GA Z
  std::vector<double> T((size_t)p * r, 0.0), H((size_t)r * r, 0.0);
  for (int i = 0; i < p; ++i)
    for (int k = 0; k < p; ++k) {
      double g = GA[(size_t)i * p + k];
      if (g == 0.0) continue;
      for (int c = 0; c < r; ++c) T[(size_t)i * r + c] += g * Z[(size_t)k * r + c];
    }
  for (int a = 0; a < r; ++a)
    for (int i = 0; i < p; ++i) {
      double z = Z[(size_t)i * r + a];
      if (z == 0.0) continue;
      for (int c = 0; c < r; ++c) H[(size_t)a * r + c] += z * T[(size_t)i * r + c];
    }
  std::vector<double> wh, Vh;
  sym_eig(r, H, wh, Vh);
  C.assign((size_t)p * m, 0.0);
  lam.assign(wh.begin(), wh.begin() + m);
  for (int i = 0; i < p; ++i)
    for (int j = 0; j < m; ++j) {
      double s = 0.0;
      for (int c = 0; c < r; ++c) s += Z[(size_t)i * r + c] * Vh[(size_t)c * r + j];
      C[(size_t)i * m + j] = s * d[i];
    }
  return true;
}

// Orthonormalization of a block U against G = U^T B U (k x k): returns T
// (k x k2, row-major) with (U T)^T B (U T) = I, dropping directions whose
// relative eigenvalue of the diagonally equilibrated G is below tau.
int svqb(int k, std::vector<double> G, double tau, std::vector<double>& T) {
  std::vector<double> d(k);
  for (int i = 0; i < k; ++i) {
    double g = G[(size_t)i * k + i];
    d[i] = g > 0 ? 1.0 / std::sqrt(g) : 0.0;
  }
  std::vector<double> H((size_t)k * k);
  for (int i = 0; i < k; ++i)
    for (int j = 0; j < k; ++j) H[(size_t)i * k + j] = 0.5 * (G[(size_t)i * k + j] + G[(size_t)j * k + i]) * d[i] * d[j];
  std::vector<double> th, V;
  sym_eig(k, H, th, V);
  double tmax = th.empty() ? 0.0 : th.back();
  std::vector<int> keep;
  for (int c = 0; c < k; ++c)
    if (th[c] > tau * tmax && th[c] > 0) keep.push_back(c);
  int k2 = (int)keep.size();
  T.assign((size_t)k * k2, 0.0);
  for (int i = 0; i < k; ++i)
    for (int c = 0; c < k2; ++c) T[(size_t)i * k2 + c] = d[i] * V[(size_t)i * k + keep[c]] / std::sqrt(th[keep[c]]);
  return k2;
}

}  // namespace

namespace tsem {

// y = A x for each of m columns (n-strided), FP64, eliminated form
static void apply_block(tsem_space* sp, double stiff, double mass, const double* X, double* Y, int m,
                        double pot = 0.0) {
  OpCoef op;
  op.stiff = stiff;
  op.mass = mass;
  op.pot = pot;
  op.constrained = 1;
  const int n = sp->d.ndof;
  for (int j = 0; j < m; ++j)
    launch_apply<double>(sp->d, op, X + (size_t)j * n, Y + (size_t)j * n, sp->yl64.as<double>(), nullptr, sp->st());
}

}  // namespace tsem

extern "C" {

int tsem_lobpcg(tsem_space* sp, int m, double* X, double* lambda, const tsem_lobpcg_opts* o,
                tsem_lobpcg_stats* stats) {
  API_BEGIN
  if (!sp || !X || !lambda || !o) throw std::invalid_argument("null argument");
  if (m < 1 || m > 32) throw std::invalid_argument("block size must be in 1..32");
  if (!(o->tol > 0)) throw std::invalid_argument("tol must be positive");
  if (o->precond != 0 && o->precond != 1 && o->precond != 4)
    throw std::invalid_argument("precond must be 0 (none), 1 (Jacobi) or 4 (p-MG)");
  if (o->criterion != 0 && o->criterion != 1) throw std::invalid_argument("criterion must be 0 or 1");
  if (!(o->stiff > 0)) throw std::invalid_argument("A = stiff S + mass M + potential M_V needs stiff > 0");
  if (o->potential != 0.0 && !sp->d.vpot) throw std::logic_error("potential coefficient set but no potential");
  hipStream_t st = sp->st();
  const int n = sp->d.ndof;
  if (n < 3 * m) throw std::invalid_argument("block too large for the space");
  auto t0 = std::chrono::steady_clock::now();
  if (o->precond == 4 && !sp->pmg) sp->pmg.reset(new Pmg(sp, st));
  const size_t nm = (size_t)n * m;
  // trial block [X | P | W] and its images under A and B: X at column 0, P
  // (np columns) at column m, W (nw columns: residuals of the pairs not yet
  // converged, preconditioned, B-orthogonalized against X and P and
  // orthonormalized) right after P
  DevBuf S, AS, BS, R, part, gram, dlam, dC, r32, z32, dscale;
  S.alloc(3 * nm * 8); AS.alloc(3 * nm * 8); BS.alloc(3 * nm * 8);
  R.alloc(nm * 8);
  part.alloc((size_t)gram_blocks(n) * 96 * 96 * 8);
  gram.alloc(96 * 96 * 8);
  dlam.alloc(32 * 8);
  dC.alloc(96 * 64 * 8);
  dscale.alloc(2 * 32 * 8);
  double* base[3] = {S.as<double>(), AS.as<double>(), BS.as<double>()};
  double* Sx = base[0];
  double* Ax = base[1];
  double* Bx = base[2];
  const uint8_t* cm = sp->d.cmask;
  int applies = 0;

  auto col = [&](int b, int c) { return base[b] + (size_t)c * n; };
  auto gram_host = [&](int p, int q, const double* U, const double* V, std::vector<double>& G) {
    G.assign((size_t)p * q, 0.0);
    if (p < 1 || q < 1) return;
    launch_gram(n, p, q, U, n, V, n, part.as<double>(), gram.as<double>(), st);
    HIPCHK(hipMemcpyAsync(G.data(), gram.p, (size_t)p * q * 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
  };
  auto upload_coef = [&](const std::vector<double>& C) {
    HIPCHK(hipMemcpyAsync(dC.p, C.data(), C.size() * 8, hipMemcpyHostToDevice, st));
    return dC.as<double>();
  };
  // columns [c0, c0 + k2) of S, AS, BS <- columns [c0, c0 + k) times Tm
  // (k x k2), in place (the combination reads a row whole before writing it)
  auto transform = [&](int c0, int k, int k2, const std::vector<double>& Tm) {
    const double* dt = upload_coef(Tm);
    for (int b = 0; b < 3; ++b) launch_combine(n, k, k2, col(b, c0), n, dt, col(b, c0), n, 1.0, 0.0, st);
  };
  // columns [w0, w0 + nw) -= columns [c0, c0 + k) Cm, Cm = ([c0..]^T B [w0..]) (k x nw)
  auto orthogonalize = [&](int c0, int k, int w0, int nw) {
    if (k < 1 || nw < 1) return;
    std::vector<double> Cm;
    gram_host(k, nw, col(0, c0), col(2, w0), Cm);
    const double* dc = upload_coef(Cm);
    for (int b = 0; b < 3; ++b) launch_combine(n, k, nw, col(b, c0), n, dc, col(b, w0), n, -1.0, 1.0, st);
  };

  // initial block: zero the eliminated rows, Rayleigh-Ritz on X alone
  HIPCHK(hipMemcpyAsync(Sx, X, nm * 8, hipMemcpyDeviceToDevice, st));
  launch_block_mask(n, m, cm, nullptr, Sx, st);
  apply_block(sp, o->stiff, o->mass, Sx, Ax, m, o->potential);
  apply_block(sp, 0.0, 1.0, Sx, Bx, m);
  applies += 2 * m;
  std::vector<double> GA, GB, C, lam, lam_last, RR, BB, Tm;
  gram_host(m, m, Sx, Ax, GA);
  gram_host(m, m, Sx, Bx, GB);
  if (!rayleigh_ritz(m, GA, GB, m, 1e-14, C, lam)) throw std::invalid_argument("initial block is rank deficient");
  transform(0, m, m, C);
  int np = 0, it = 0;
  double maxres = 0.0, maxdl = 0.0;
  bool conv = false;
  static const bool trace = getenv("TSEM_LOBPCG_TRACE") != nullptr;
  for (; it < o->max_iter; ++it) {
    // residuals and their relative norms
    HIPCHK(hipMemcpyAsync(dlam.p, lam.data(), (size_t)m * 8, hipMemcpyHostToDevice, st));
    launch_block_resid(n, m, Ax, Bx, dlam.as<double>(), cm, R.as<double>(), st);
    gram_host(m, m, R.as<double>(), R.as<double>(), RR);
    gram_host(m, m, Bx, Bx, BB);
    maxres = 0.0;
    std::vector<double> res(m);
    for (int j = 0; j < m; ++j) {
      res[j] = std::sqrt(std::max(0.0, RR[(size_t)j * m + j])) /
               (std::fabs(lam[j]) * std::sqrt(std::max(1e-300, BB[(size_t)j * m + j])));
      maxres = std::max(maxres, res[j]);
    }
    maxdl = 0.0;
    if (!lam_last.empty())
      for (int j = 0; j < m; ++j) maxdl = std::max(maxdl, std::fabs(lam[j] - lam_last[j]));
    if (trace) fprintf(stderr, "[tsem lobpcg] it %d maxres %.3e maxdl %.3e np %d lam0 %.12e\n", it, maxres, maxdl, np,
                       lam[0]);
    if (o->criterion == 1 ? maxres <= o->tol : (!lam_last.empty() && maxdl < o->tol)) {
      conv = true;
      break;
    }
    lam_last = lam;
    // W: residuals of the pairs not converged yet (soft locking)
    const int w0 = m + np;
    int nw = 0;
    std::vector<double> wnorm;
    for (int j = 0; j < m; ++j) {
      if (o->criterion == 1 && res[j] <= o->tol) continue;
      HIPCHK(hipMemcpyAsync(col(0, w0 + nw), R.as<double>() + (size_t)j * n, (size_t)n * 8, hipMemcpyDeviceToDevice,
                            st));
      wnorm.push_back(std::sqrt(std::max(0.0, RR[(size_t)j * m + j])));
      ++nw;
    }
    double* Wp = col(0, w0);
    if (o->precond == 1) {
      launch_block_mask(n, nw, cm, sp->d.invdiag, Wp, st);
    } else if (o->precond == 4 && nw > 1 && block_vcycle_enabled() && sp->pmg->block_supported()) {
      // every column r_k / |r_k| through one block V-cycle, scaled back
      r32.alloc((size_t)n * nw * 4);
      z32.alloc((size_t)n * nw * 4);
      std::vector<double> sc(2 * (size_t)nw);
      for (int k = 0; k < nw; ++k) {
        sc[k] = wnorm[k] > 0 ? 1.0 / wnorm[k] : 0.0;
        sc[nw + k] = wnorm[k] > 0 ? wnorm[k] : 0.0;
      }
      HIPCHK(hipMemcpyAsync(dscale.p, sc.data(), sc.size() * 8, hipMemcpyHostToDevice, st));
      launch_cast_cols(n, nw, Wp, dscale.as<double>(), r32.as<float>(), st);
      sp->pmg->vcycle_replay<float>(r32.as<float>(), z32.as<float>(), o->smooth_steps, st, nw);
      launch_widen_cols(n, nw, z32.as<float>(), dscale.as<double>() + nw, Wp, st);
      launch_block_mask(n, nw, cm, nullptr, Wp, st);
    } else if (o->precond == 4) {
      r32.alloc((size_t)n * 4);
      z32.alloc((size_t)n * 4);
      for (int k = 0; k < nw; ++k) {
        double nr = wnorm[k];
        if (!(nr > 0)) continue;
        double* w = Wp + (size_t)k * n;
        launch_cast_f64_f32(n, w, 1.0 / nr, r32.as<float>(), st);
        sp->pmg->vcycle_replay<float>(r32.as<float>(), z32.as<float>(), o->smooth_steps, st);
        launch_cast_f32_f64(n, z32.as<float>(), nr, w, st);
      }
      launch_block_mask(n, nw, cm, nullptr, Wp, st);
    }
    apply_block(sp, o->stiff, o->mass, Wp, col(1, w0), nw, o->potential);
    apply_block(sp, 0.0, 1.0, Wp, col(2, w0), nw);
    applies += 2 * nw;
    // B-orthogonalize W against X and P, then orthonormalize it
    orthogonalize(0, m, w0, nw);
    orthogonalize(m, np, w0, nw);
    if (nw > 0) {
      gram_host(nw, nw, Wp, col(2, w0), GB);
      int nw2 = svqb(nw, GB, 1e-12, Tm);
      transform(w0, nw, nw2, Tm);
      nw = nw2;
    }
    // Rayleigh-Ritz on [X P W]
    int p = m + np + nw;
    gram_host(p, p, Sx, Ax, GA);
    gram_host(p, p, Sx, Bx, GB);
    if (!rayleigh_ritz(p, GA, GB, m, 1e-13, C, lam)) throw std::runtime_error("LOBPCG trial space collapsed");
    // X <- [X P W] C and P <- [P W] C[m:] as one in-place combination with
    // the coefficients [C | (0; C[m:])] (p x 2m), then orthonormalize P
    const int m2 = p > m ? 2 * m : m;
    std::vector<double> C2((size_t)p * m2, 0.0);
    for (int i = 0; i < p; ++i)
      for (int j = 0; j < m; ++j) {
        C2[(size_t)i * m2 + j] = C[(size_t)i * m + j];
        if (p > m && i >= m) C2[(size_t)i * m2 + m + j] = C[(size_t)i * m + j];
      }
    const double* dc = upload_coef(C2);
    for (int b = 0; b < 3; ++b) {
      if (m2 > m && m % 16 == 0)
        launch_combine_split(n, p, m2, base[b], n, dc, base[b], n, m, m, st);
      else
        launch_combine(n, p, m2, base[b], n, dc, base[b], n, 1.0, 0.0, st);
    }
    np = 0;
    if (p > m) {
      gram_host(m, m, col(0, m), col(2, m), GB);
      np = svqb(m, GB, 1e-12, Tm);
      transform(m, m, np, Tm);
    }
  }
  HIPCHK(hipMemcpyAsync(X, Sx, nm * 8, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
  for (int j = 0; j < m; ++j) lambda[j] = lam[j];
  if (stats) {
    stats->iterations = it;
    stats->max_resid = maxres;
    stats->max_dlambda = maxdl;
    stats->applies = applies;
    stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  if (!conv) { g_err = "LOBPCG did not reach the tolerance"; return TSEM_ENOCONV; }
  API_END
}

int tsem_inverse_power(tsem_space* sp, double* x, double tol, int max_iter, const tsem_pcg_opts* inner,
                       double* lambda, int* iterations) {
  API_BEGIN
  if (!sp || !x || !inner || !lambda) throw std::invalid_argument("null argument");
  if (!(tol > 0)) throw std::invalid_argument("tol must be positive");
  hipStream_t st = sp->st();
  const int n = sp->d.ndof;
  DevBuf rhs, ax, part, g;
  rhs.alloc((size_t)n * 8); ax.alloc((size_t)n * 8);
  part.alloc((size_t)gram_blocks(n) * 8); g.alloc(8);
  auto dot = [&](const double* a, const double* b) {
    double h = 0;
    launch_gram(n, 1, 1, a, n, b, n, part.as<double>(), g.as<double>(), st);
    HIPCHK(hipMemcpyAsync(&h, g.p, 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return h;
  };
  launch_block_mask(n, 1, sp->d.cmask, nullptr, x, st);
  apply_block(sp, 0.0, 1.0, x, rhs.as<double>(), 1);
  double e_last = 0, e_now = 0;
  int it = 0;
  bool conv = false;
  for (; it < max_iter; ++it) {
    tsem_pcg_stats s;
    int rc = tsem_pcg(sp, rhs.as<double>(), x, inner, &s);
    if (rc != TSEM_OK && rc != TSEM_ENOCONV) return rc;
    apply_block(sp, 0.0, 1.0, x, rhs.as<double>(), 1);
    double c = std::sqrt(dot(rhs.as<double>(), x));
    launch_axpby<double>(n, 1.0 / c, x, 0.0, x, st);
    launch_axpby<double>(n, 1.0 / c, rhs.as<double>(), 0.0, rhs.as<double>(), st);
    apply_block(sp, 1.0, 0.0, x, ax.as<double>(), 1);
    e_now = dot(ax.as<double>(), x);
    if (it > 0 && std::fabs(e_now - e_last) < tol) { conv = true; ++it; break; }
    e_last = e_now;
  }
  *lambda = e_now;
  if (iterations) *iterations = it;
  if (!conv) { g_err = "inverse power iteration did not reach the tolerance"; return TSEM_ENOCONV; }
  API_END
}

}  // extern "C"
