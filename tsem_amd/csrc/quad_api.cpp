// C ABI of the quadrature-point work (include/tsem.h): point values, density,
// integrals and moments, L2 error, LDA exchange-correlation, and the
// potential that weights the M_V operator.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <limits>
#include <vector>

#include "handle.hpp"

using namespace tsem;

namespace {

// basis values and weights of the space's 3-D quadrature on the device
void ensure_quad(tsem_space* sp) {
  if (sp->qphi.p) return;
  const Space& S = sp->h;
  const int K = S.K, nq = sp->nq3;
  std::vector<double> phi((size_t)nq * K), phiT((size_t)K * nq);
  std::vector<long double> val(K);
  for (int q = 0; q < nq; ++q) {
    S.ref.eval(sp->p3[3 * q], sp->p3[3 * q + 1], sp->p3[3 * q + 2], val.data(), nullptr);
    for (int i = 0; i < K; ++i) {
      phi[(size_t)q * K + i] = (double)val[i];
      phiT[(size_t)i * nq + q] = (double)val[i];
    }
  }
  hipStream_t st = sp->st();
  upload(sp->qphi, phi, st);
  upload(sp->qphiT, phiT, st);
  upload(sp->qw, sp->w3, st);
  DevSpace& d = sp->d;
  d.nq = nq;
  d.phi = sp->qphi.as<double>();
  d.phiT = sp->qphiT.as<double>();
  d.wq = sp->qw.as<double>();
  d.xe = sp->xe.as<double>();
  d.pts = sp->pts3.as<double>();
  HIPCHK(hipStreamSynchronize(st));
}

// reference basis gradients at the 3-D points, built on first use
void ensure_quad_grad(tsem_space* sp) {
  ensure_quad(sp);
  if (sp->qdphi.p) return;
  const Space& S = sp->h;
  const int K = S.K, nq = sp->nq3;
  std::vector<double> dphi((size_t)nq * K * 3);
  std::vector<long double> val(K), grad((size_t)K * 3);
  for (int q = 0; q < nq; ++q) {
    S.ref.eval(sp->p3[3 * q], sp->p3[3 * q + 1], sp->p3[3 * q + 2], val.data(), grad.data());
    for (int i = 0; i < 3 * K; ++i) dphi[(size_t)q * K * 3 + i] = (double)grad[i];
  }
  upload(sp->qdphi, dphi, sp->st());
  sp->d.dphi = sp->qdphi.as<double>();
  HIPCHK(hipStreamSynchronize(sp->st()));
}

// bucket the element bounding boxes on a uniform grid of about one element
// per cell over the mesh's bounding box
void ensure_locator(tsem_space* sp) {
  auto& L = sp->loc;
  if (L.valid) return;
  const Space& S = sp->h;
  const int nt = S.topo.nt;
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int v = 0; v < S.topo.nv; ++v)
    for (int d = 0; d < 3; ++d) {
      lo[d] = std::min(lo[d], S.xyz[3 * (size_t)v + d]);
      hi[d] = std::max(hi[d], S.xyz[3 * (size_t)v + d]);
    }
  const int nc = std::max(1, (int)std::ceil(std::cbrt((double)nt)));
  for (int d = 0; d < 3; ++d) {
    L.n[d] = nc;
    L.lo[d] = lo[d];
    L.h[d] = hi[d] > lo[d] ? (hi[d] - lo[d]) / nc : 1.0;
  }
  auto cell = [&](double c, int d) { return std::min(L.n[d] - 1, std::max(0, (int)std::floor((c - L.lo[d]) / L.h[d]))); };
  std::vector<std::array<int, 6>> box(nt);
  for (int t = 0; t < nt; ++t) {
    double a[3] = {1e300, 1e300, 1e300}, b[3] = {-1e300, -1e300, -1e300};
    for (int k = 0; k < 4; ++k)
      for (int d = 0; d < 3; ++d) {
        const double c = S.xyz[3 * (size_t)S.tet[4 * (size_t)t + k] + d];
        a[d] = std::min(a[d], c);
        b[d] = std::max(b[d], c);
      }
    for (int d = 0; d < 3; ++d) { box[t][d] = cell(a[d], d); box[t][3 + d] = cell(b[d], d); }
  }
  const size_t ncell = (size_t)L.n[0] * L.n[1] * L.n[2];
  L.ptr.assign(ncell + 1, 0);
  for (int pass = 0; pass < 2; ++pass) {
    std::vector<int32_t> fill;
    if (pass == 1) {
      for (size_t c = 0; c < ncell; ++c) L.ptr[c + 1] += L.ptr[c];
      L.elems.assign(L.ptr[ncell], 0);
      fill.assign(L.ptr.begin(), L.ptr.end() - 1);
    }
    for (int t = 0; t < nt; ++t)
      for (int i = box[t][0]; i <= box[t][3]; ++i)
        for (int j = box[t][1]; j <= box[t][4]; ++j)
          for (int k = box[t][2]; k <= box[t][5]; ++k) {
            const size_t c = ((size_t)i * L.n[1] + j) * L.n[2] + k;
            if (pass == 0) ++L.ptr[c + 1];
            else L.elems[fill[c]++] = t;
          }
  }
  L.valid = true;
}

// reference coordinates of p in element t (p = P0 + J r, J = [P1-P0, P2-P0,
// P3-P0]) when its barycentric coordinates are all >= -tol
bool in_element(const Space& S, int t, const double* p, double tol, double* r) {
  const double* P[4];
  for (int k = 0; k < 4; ++k) P[k] = &S.xyz[3 * (size_t)S.tet[4 * (size_t)t + k]];
  double J[3][3], q[3];
  for (int d = 0; d < 3; ++d) {
    for (int c = 0; c < 3; ++c) J[d][c] = P[c + 1][d] - P[0][d];
    q[d] = p[d] - P[0][d];
  }
  const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                     J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
  if (det == 0) return false;
  // Cramer's rule
  for (int c = 0; c < 3; ++c) {
    double A[3][3];
    for (int d = 0; d < 3; ++d)
      for (int e = 0; e < 3; ++e) A[d][e] = (e == c) ? q[d] : J[d][e];
    r[c] = (A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) - A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
            A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0])) / det;
  }
  return r[0] >= -tol && r[1] >= -tol && r[2] >= -tol && 1 - r[0] - r[1] - r[2] >= -tol;
}

}  // namespace

extern "C" {

int tsem_qp_values(tsem_space* sp, const double* x, int e0, int e1, double* out) {
  API_BEGIN
  if (!sp || !x || !out) throw std::invalid_argument("null argument");
  if (e0 < 0 || e1 > sp->d.nt || e0 > e1) throw std::invalid_argument("element range out of bounds");
  ensure_quad(sp);
  sp->qxl.alloc((size_t)sp->d.nt * sp->d.K * 8);
  launch_gather(sp->d, x, sp->qxl.as<double>(), 0, sp->st());
  launch_qp_values(sp->d, e0, e1, sp->qxl.as<double>(), out, sp->st());
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(sp->st()));
  API_END
}

int tsem_density(tsem_space* sp, int nk, const double* psi, const double* occ, double* rho) {
  API_BEGIN
  if (!sp || !psi || !occ || !rho) throw std::invalid_argument("null argument");
  if (nk < 1 || (size_t)nk * sp->d.K * 8 > 64 * 1024) throw std::invalid_argument("number of states out of range");
  ensure_quad(sp);
  const DevSpace& d = sp->d;
  sp->qxl.alloc((size_t)nk * d.nt * d.K * 8);
  for (int k = 0; k < nk; ++k)
    launch_gather(d, psi + (size_t)k * d.ndof, sp->qxl.as<double>() + (size_t)k * d.nt * d.K, 0, sp->st());
  DevBuf docc;
  upload(docc, std::vector<double>(occ, occ + nk), sp->st());
  launch_density(d, nk, sp->qxl.as<double>(), docc.as<double>(), rho, sp->st());
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(sp->st()));
  API_END
}

int tsem_integrate(tsem_space* sp, const double* f, const double* g, int nmom, double* out) {
  API_BEGIN
  if (!sp || !f || !out) throw std::invalid_argument("null argument");
  if (nmom != 1 && nmom != 10) throw std::invalid_argument("nmom must be 1 or 10");
  ensure_quad(sp);
  sp->qpart.alloc((size_t)10 * integrate_blocks(sp->d) * 8);
  sp->qout.alloc(10 * 8);
  launch_integrate(sp->d, f, g, nmom, sp->qpart.as<double>(), sp->qout.as<double>(), sp->st());
  HIPCHK(hipMemcpyAsync(out, sp->qout.p, (size_t)nmom * 8, hipMemcpyDeviceToHost, sp->st()));
  HIPCHK(hipStreamSynchronize(sp->st()));
  API_END
}

int tsem_l2_error_chunk(tsem_space* sp, const double* x, int e0, int e1, const double* u, double* sum) {
  API_BEGIN
  if (!sp || !x || !u || !sum) throw std::invalid_argument("null argument");
  if (e0 < 0 || e1 > sp->d.nt || e0 > e1) throw std::invalid_argument("element range out of bounds");
  ensure_quad(sp);
  sp->qxl.alloc((size_t)sp->d.nt * sp->d.K * 8);
  sp->qpart.alloc((size_t)10 * integrate_blocks(sp->d) * 8);
  sp->qout.alloc(10 * 8);
  launch_gather(sp->d, x, sp->qxl.as<double>(), 0, sp->st());
  double h = 0.0;
  if (e1 > e0) {
    launch_l2_chunk(sp->d, e0, e1, sp->qxl.as<double>(), u, sp->qpart.as<double>(), sp->qout.as<double>(), sp->st());
    HIPCHK(hipMemcpyAsync(&h, sp->qout.p, 8, hipMemcpyDeviceToHost, sp->st()));
  }
  HIPCHK(hipStreamSynchronize(sp->st()));
  *sum = h;
  API_END
}

int tsem_h1_error_chunk(tsem_space* sp, const double* x, int e0, int e1, const double* g, double* sum) {
  API_BEGIN
  if (!sp || !x || !g || !sum) throw std::invalid_argument("null argument");
  if (e0 < 0 || e1 > sp->d.nt || e0 > e1) throw std::invalid_argument("element range out of bounds");
  ensure_quad_grad(sp);
  sp->qxl.alloc((size_t)sp->d.nt * sp->d.K * 8);
  sp->qpart.alloc((size_t)10 * integrate_blocks(sp->d) * 8);
  sp->qout.alloc(10 * 8);
  launch_gather(sp->d, x, sp->qxl.as<double>(), 0, sp->st());
  double h = 0.0;
  if (e1 > e0) {
    launch_h1_chunk(sp->d, e0, e1, sp->qxl.as<double>(), g, sp->qpart.as<double>(), sp->qout.as<double>(), sp->st());
    HIPCHK(hipMemcpyAsync(&h, sp->qout.p, 8, hipMemcpyDeviceToHost, sp->st()));
  }
  HIPCHK(hipStreamSynchronize(sp->st()));
  *sum = h;
  API_END
}

int tsem_lda_pz(tsem_space* sp, int64_t n, const double* rho, double* vxc, double* exc) {
  API_BEGIN
  if (!sp || !rho) throw std::invalid_argument("null argument");
  launch_lda_pz((long)n, rho, vxc, exc, sp->st());
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(sp->st()));
  API_END
}

int tsem_set_potential(tsem_space* sp, const double* v) {
  API_BEGIN
  if (!sp) throw std::invalid_argument("null space");
  ensure_quad(sp);
  if (!v) {
    sp->d.vpot = nullptr;
  } else {
    size_t bytes = (size_t)sp->d.nt * sp->d.nq * 8;
    sp->vpot.alloc(bytes);
    HIPCHK(hipMemcpyAsync(sp->vpot.p, v, bytes, hipMemcpyDeviceToDevice, sp->st()));
    sp->d.vpot = sp->vpot.as<double>();
  }
  HIPCHK(hipStreamSynchronize(sp->st()));
  API_END
}

int tsem_apply_potential(tsem_space* sp, double stiff, double mass, double potential, int constrained,
                         const double* x, double* y) {
  API_BEGIN
  if (!sp || !x || !y) throw std::invalid_argument("null argument");
  if (potential != 0.0 && !sp->d.vpot) throw std::logic_error("no potential set");
  OpCoef op;
  op.stiff = stiff; op.mass = mass; op.pot = potential; op.constrained = constrained;
  launch_apply<double>(sp->d, op, x, y, sp->yl64.as<double>(), nullptr, sp->st());
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(sp->st()));
  API_END
}

int tsem_space_set_geometry(tsem_space* sp, const double* xyz) {
  API_BEGIN
  if (!sp || !xyz) throw std::invalid_argument("null argument");
  Space& S = sp->h;
  std::vector<double> keep(S.xyz);
  S.set_geometry(xyz);
  for (int t = 0; t < S.topo.nt; ++t)
    if (!(S.vol[t] > 0)) {
      S.set_geometry(keep.data());
      throw std::invalid_argument("degenerate tetrahedron");
    }
  hipStream_t st = sp->st();
  DevSpace& d = sp->d;
  upload(sp->ghat, S.ghat, st);
  upload(sp->ghat32, std::vector<float>(S.ghat.begin(), S.ghat.end()), st);
  upload(sp->sixv, S.sixv, st);
  upload(sp->sixv32, std::vector<float>(S.sixv.begin(), S.sixv.end()), st);
  std::vector<double> xe((size_t)S.topo.nt * 12);
  for (int e = 0; e < S.topo.nt; ++e)
    for (int k = 0; k < 4; ++k)
      for (int c = 0; c < 3; ++c) xe[(size_t)e * 12 + 3 * k + c] = S.xyz[3 * (size_t)S.tet[4 * e + k] + c];
  upload(sp->xe, xe, st);
  launch_diagonal(d, sp->grad2.as<double>(), sp->diag_d.as<double>(), sp->invdiag.as<double>(),
                  sp->invdiag32.as<float>(), st);
  HIPCHK(hipGetLastError());
  sp->diag.resize(S.ndof);
  HIPCHK(hipMemcpyAsync(sp->diag.data(), sp->diag_d.p, (size_t)S.ndof * 8, hipMemcpyDeviceToHost, st));
  // a new geometry starts without boundary conditions and potential
  std::fill(sp->cmask.begin(), sp->cmask.end(), (uint8_t)0);
  upload_mask(sp, st);
  d.vpot = nullptr;
  sp->drop_pmg();
  sp->loc.valid = false;
  HIPCHK(hipStreamSynchronize(st));
  API_END
}

}  // extern "C"

extern "C" int tsem_point_values(tsem_space* sp, const double* x, int64_t n, const double* pts, double* out,
                                 int32_t* elem) {
  API_BEGIN
  if (!sp || !x || (n > 0 && (!pts || !out))) throw std::invalid_argument("null argument");
  if (n < 0) throw std::invalid_argument("negative point count");
  if (n == 0) return TSEM_OK;
  const Space& S = sp->h;
  const int K = S.K;
  ensure_locator(sp);
  const auto& L = sp->loc;
  // host: the element holding each point (lowest index among the candidates
  // of its cell), its reference coordinates and the basis there
  std::vector<int32_t> ep(n, -1);
  std::vector<double> phi((size_t)n * K, 0.0);
  std::vector<long double> val(K);
  const double tol = 1e-12;
  for (int64_t p = 0; p < n; ++p) {
    const double* q = pts + 3 * p;
    int c[3];
    bool inside = true;
    for (int d = 0; d < 3; ++d) {
      const double f = std::floor((q[d] - L.lo[d]) / L.h[d]);
      if (f < -1 || f > L.n[d]) inside = false;
      c[d] = std::min(L.n[d] - 1, std::max(0, (int)f));
    }
    if (!inside) continue;
    const size_t cell = ((size_t)c[0] * L.n[1] + c[1]) * L.n[2] + c[2];
    double r[3];
    for (int32_t k = L.ptr[cell]; k < L.ptr[cell + 1]; ++k) {
      const int t = L.elems[k];
      if (in_element(S, t, q, tol, r)) {
        ep[p] = t;
        S.ref.eval(r[0], r[1], r[2], val.data(), nullptr);
        for (int i = 0; i < K; ++i) phi[(size_t)p * K + i] = (double)val[i];
        break;
      }
    }
  }
  // device: element coefficients, then one fixed-order sum per point
  hipStream_t st = sp->st();
  sp->qxl.alloc((size_t)S.topo.nt * K * 8);
  launch_gather(sp->d, x, sp->qxl.as<double>(), 0, st);
  DevBuf dep, dphi, dout;
  upload(dep, ep, st);
  upload(dphi, phi, st);
  dout.alloc((size_t)n * 8);
  launch_point_values(K, n, dep.as<int32_t>(), dphi.as<double>(), sp->qxl.as<double>(), dout.as<double>(), st);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, dout.p, (size_t)n * 8, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (elem) std::copy(ep.begin(), ep.end(), elem);
  API_END
}
