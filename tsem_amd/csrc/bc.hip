// Boundary conditions on the device: Neumann load and Dirichlet modal
// interpolation with elimination (TSEM::impose_boundary_condition,
// src/Discretize.cpp:1778-1893 and 2010-2063).  The load vector stays on the
// device; only the caller's boundary data (host arrays, given per marked
// geometry at tsem_bc_points) are copied in.
//
// Dirichlet: vertices take their data; an edge's modes are the 1-D
// projection of its data minus the linear interpolant of its end values onto
// J^{1,1}_{l-2}; a face's modes are the 2-D projection of its data minus the
// vertex and edge traces (src/BuildSpace.cpp:95-147 bases).  Every geometry
// writes only its own rows, one thread per (geometry, mode); vertex and edge
// values are read back as b / diag, as the reference reads rhs(row) /
// diag_element (src/Discretize.cpp:1837-1880).
// Neumann: per marked face, the load of each of its vertex, edge and face
// modes by the face's 2-D rule (parallel over faces and modes), then added
// into b row by row in face order (rows shared by faces get their terms in
// one fixed order).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "handle.hpp"

namespace tsem {
namespace {
constexpr int BLK = 256;
inline unsigned nblk(long n) { return (unsigned)((n + BLK - 1) / BLK); }

// standard Jacobi P_n^{(a,b)}(x), a, b > -1 (three-term recurrence)
double jacobi_p(double a, double b, int n, double x) {
  if (n == 0) return 1.0;
  double p0 = 1.0, p1 = (a - b) / 2 + (a + b + 2) / 2 * x;
  for (int k = 2; k <= n; ++k) {
    const double c = 2 * k + a + b;
    const double A1 = 2 * k * (k + a + b) * (c - 2);
    const double A2 = (c - 1) * (a * a - b * b);
    const double A3 = (c - 1) * c * (c - 2);
    const double A4 = 2 * (k + a - 1) * (k + b - 1) * c;
    const double p2 = ((A2 + A3 * x) * p1 - A4 * p0) / A1;
    p0 = p1;
    p1 = p2;
  }
  return p1;
}

// geometries of dimension dim carrying mark `flag`, in index order
std::vector<int> marked_geos(const Space& S, int dim, int flag) {
  if (S.mark[dim].empty()) throw std::logic_error("boundary marks not set");
  std::vector<int> out;
  for (size_t g = 0; g < S.mark[dim].size(); ++g)
    if (S.mark[dim][g] == flag) out.push_back((int)g);
  return out;
}

// offsets of the boundary tables in tsem_space::bc_tab (doubles)
enum { T_T1, T_W1, T_P2, T_W2, T_IE1, T_IF2, T_EA, T_EB, T_BF, T_COUNT };

__global__ void k_dir_vertex(int n, const int32_t* __restrict__ row, const double* __restrict__ u,
                             const double* __restrict__ dg, double* __restrict__ b, double* __restrict__ vb) {
  const int j = blockIdx.x * BLK + threadIdx.x;
  if (j >= n) return;
  const int r = row[j];
  b[r] = dg[r] * u[j];
  vb[r] = u[j];
}

// marked edge j, mode l = 2 + m: row0[j] + m; end vertices' rows vs[j], ve[j]
__global__ void k_dir_edge(int ne, int nde, int q1, const int32_t* __restrict__ row0, const int32_t* __restrict__ vs,
                           const int32_t* __restrict__ ve, const double* __restrict__ ue, const double* __restrict__ tab,
                           const long* __restrict__ off, const double* __restrict__ dg, double* __restrict__ b,
                           double* __restrict__ vb) {
#pragma clang fp contract(off)
  const long s = (long)blockIdx.x * BLK + threadIdx.x;
  if (s >= (long)ne * nde) return;
  const int j = (int)(s / nde), m = (int)(s % nde), l = m + 2;
  const double* t1 = tab + off[T_T1];
  const double* w1 = tab + off[T_W1];
  const double* ie1 = tab + off[T_IE1];
  const double cs = b[vs[j]] / dg[vs[j]], ce = b[ve[j]] / dg[ve[j]];
  double c = 0;
  for (int k = 0; k < q1; ++k)
    c += w1[k] * (ue[(size_t)j * q1 + k] - cs * (1 - t1[k]) - ce * t1[k]) * ie1[(size_t)k * nde + m];
  const double val = c * -l * (2 * l - 1) / (2 * (l - 1));
  const int r = row0[j] + m;
  b[r] = dg[r] * val;
  vb[r] = val;
}

// marked face j, face mode id (l1 + l2 = s, id = (s-3)(s-2)/2 + l2 - 1):
// rows frow[j] + id; fv[3j..]: vertex rows, fe[3j..]: edge DOF starts,
// frev[j]: bit ie set when edge ie runs against the face
__global__ void k_dir_face(int nf, int ndf, int nde, int q2, const int32_t* __restrict__ frow,
                           const int32_t* __restrict__ fv, const int32_t* __restrict__ fe,
                           const int32_t* __restrict__ frev, const double* __restrict__ uf,
                           const double* __restrict__ tab, const long* __restrict__ off,
                           const double* __restrict__ dg, double* __restrict__ b, double* __restrict__ vb) {
#pragma clang fp contract(off)
  constexpr int NDE_MAX = 11;  // orders <= 12
  const long s = (long)blockIdx.x * BLK + threadIdx.x;
  if (s >= (long)nf * ndf) return;
  const int j = (int)(s / ndf), id = (int)(s % ndf);
  int sd = 3;
  while ((sd - 2) * (sd - 1) / 2 <= id) ++sd;  // l1 + l2
  const int l2 = id - (sd - 3) * (sd - 2) / 2 + 1, l1 = sd - l2;
  const double* p2 = tab + off[T_P2];
  const double* w2 = tab + off[T_W2];
  const double* if2 = tab + off[T_IF2];
  const double* ea = tab + off[T_EA];
  const double* eb = tab + off[T_EB];
  double cvt[3];
  for (int v = 0; v < 3; ++v) { const int r = fv[3 * j + v]; cvt[v] = b[r] / dg[r]; }
  double ved[3][NDE_MAX];
  for (int ie = 0; ie < 3; ++ie) {
    const int es = fe[3 * j + ie];
    const bool rev = (frev[j] >> ie) & 1;
    for (int m = 0; m < nde; ++m) {
      const double v = b[es + m] / dg[es + m];
      ved[ie][m] = (rev && (m + 2) % 2 == 1) ? -v : v;
    }
  }
  double c = 0;
  for (int k = 0; k < q2; ++k) {
    const double x = p2[3 * k], y = p2[3 * k + 1], r = p2[3 * k + 2];
    const double bq[3] = {x, y, r};
    double cp = uf[(size_t)j * q2 + k];
    for (int v = 0; v < 3; ++v) cp -= cvt[v] * bq[(v + 2) % 3];
    for (int m = 0; m < nde; ++m)
      cp += 2 * (x * y * ved[0][m] * eb[(size_t)k * nde + m] + y * r * ved[1][m] * eb[(size_t)k * nde + m] +
                 x * r * ved[2][m] * ea[(size_t)k * nde + m]);
    c += w2[k] * cp * if2[(size_t)k * ndf + id];
  }
  const double val = c * -l1 * (2 * l1 - 1) * (2 * l1 + 2 * l2 - 1) / (std::pow(2.0, l1) * (l1 - 1));
  const int rr = frow[j] + id;
  b[rr] = dg[rr] * val;
  vb[rr] = val;
}

// b -= w on the rows flagged free (neither imposed now nor eliminated before)
__global__ void k_dir_fix(int n, const uint8_t* __restrict__ freerow, const double* __restrict__ w,
                          double* __restrict__ b) {
  const int i = blockIdx.x * BLK + threadIdx.x;
  if (i < n && freerow[i]) b[i] -= w[i];
}

// Neumann load of marked face j, local slot t: 0..2 vertices, then edge ie
// modes (3 + ie nde + m), then face modes (3 + 3 nde + dd)
__global__ void k_neu_local(int nf, int nloc, int nde, int ndf, int q2, const double* __restrict__ area,
                            const double* __restrict__ gq, const double* __restrict__ tab,
                            const long* __restrict__ off, double* __restrict__ out) {
#pragma clang fp contract(off)
  const long s = (long)blockIdx.x * BLK + threadIdx.x;
  if (s >= (long)nf * nloc) return;
  const int j = (int)(s / nloc), t = (int)(s % nloc);
  const double* p2 = tab + off[T_P2];
  const double* w2 = tab + off[T_W2];
  const double* ea = tab + off[T_EA];
  const double* eb = tab + off[T_EB];
  const double* bf = tab + off[T_BF];
  double c = 0;
  for (int k = 0; k < q2; ++k) {
    const double x = p2[3 * k], y = p2[3 * k + 1], r = p2[3 * k + 2];
    const double wg = w2[k] * gq[(size_t)j * q2 + k] * area[j];
    double phi;
    if (t < 3) {
      const double bq[3] = {x, y, r};
      phi = bq[(t + 2) % 3];
    } else if (t < 3 + 3 * nde) {
      const int ie = (t - 3) / nde, m = (t - 3) % nde;
      phi = ie == 0 ? -2 * x * y * eb[(size_t)k * nde + m]
                    : ie == 1 ? -2 * y * r * eb[(size_t)k * nde + m] : -2 * x * r * ea[(size_t)k * nde + m];
    } else {
      phi = bf[(size_t)k * ndf + (t - 3 - 3 * nde)];
    }
    c += wg * phi;
  }
  out[s] = c;
}

// b[row[i]] += sum of sgn * contrib[slot] over the row's entries, in order
__global__ void k_neu_add(int nrow, const int32_t* __restrict__ row, const int32_t* __restrict__ ptr,
                          const int32_t* __restrict__ slot, const double* __restrict__ contrib,
                          double* __restrict__ b) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * BLK + threadIdx.x;
  if (i >= nrow) return;
  double acc = b[row[i]];
  for (int k = ptr[i]; k < ptr[i + 1]; ++k) {
    const int sl = slot[k];
    const double v = contrib[sl & 0x7fffffff];
    acc += (sl < 0) ? -v : v;
  }
  b[row[i]] = acc;
}

// quadrature and modal bases of the boundary work, built once per space
void bc_tables(tsem_space* sp) {
  if (sp->bc_off.size() == T_COUNT + 1) return;
  const Space& S = sp->h;
  const int M = S.M, q1 = sp->nq1, q2 = sp->nq2, nde = S.ndg[1], ndf = S.ndg[2];
  std::vector<double> t1(sp->t1), w1(sp->w1), p2(sp->p2), w2(sp->w2);
  std::vector<double> ie1((size_t)q1 * std::max(nde, 1)), if2((size_t)q2 * std::max(ndf, 1)),
      ea((size_t)q2 * std::max(nde, 1)), eb((size_t)q2 * std::max(nde, 1)), bf((size_t)q2 * std::max(ndf, 1));
  for (int k = 0; k < q1; ++k)
    for (int l = 2; l <= M; ++l) ie1[(size_t)k * nde + l - 2] = jacobi_p(1, 1, l - 2, 2 * sp->t1[k] - 1);
  std::vector<long double> val(S.K);
  for (int k = 0; k < q2; ++k) {
    const double x = sp->p2[3 * k], y = sp->p2[3 * k + 1];
    const double xi = 2 * x / (1 - y) - 1, eta = 2 * y - 1;
    for (int l1 = 2; l1 <= M; ++l1)
      for (int l2 = 1; l1 + l2 <= M; ++l2)
        if2[(size_t)k * ndf + (l1 + l2 - 3) * (l1 + l2 - 2) / 2 + l2 - 1] =
            std::pow(1 - eta, l1 - 2) * jacobi_p(1, 1, l1 - 2, xi) * jacobi_p(2 * l1 - 1, 1, l2 - 1, eta);
    for (int l = 2; l <= M; ++l) {
      ea[(size_t)k * nde + l - 2] = std::pow(1 - y, l - 2) * jacobi_p(1, 1, l - 2, xi);
      eb[(size_t)k * nde + l - 2] = jacobi_p(1, 1, l - 2, eta);
    }
    // face modes of the element basis on the face z = 0 (TSEM basis_value[1])
    S.ref.eval(x, y, 0.0L, val.data(), nullptr);
    for (int l1 = 2; l1 <= M; ++l1)
      for (int l2 = 1; l1 + l2 <= M; ++l2)
        bf[(size_t)k * ndf + (l1 + l2 - 3) * (l1 + l2 - 2) / 2 + l2 - 1] = (double)val[S.ref.idx.number(l1, l2, 0)];
  }
  const std::vector<double>* parts[T_COUNT] = {&t1, &w1, &p2, &w2, &ie1, &if2, &ea, &eb, &bf};
  std::vector<long> off(T_COUNT + 1, 0);
  for (int k = 0; k < T_COUNT; ++k) off[k + 1] = off[k] + (long)parts[k]->size();
  std::vector<double> all(off[T_COUNT]);
  for (int k = 0; k < T_COUNT; ++k) std::copy(parts[k]->begin(), parts[k]->end(), all.begin() + off[k]);
  upload(sp->bc_tab, all, sp->st());
  upload(sp->bc_offd, off, sp->st());
  HIPCHK(hipStreamSynchronize(sp->st()));
  sp->bc_off = off;
}

template <typename T>
void put(DevBuf& d, const std::vector<T>& v, hipStream_t st) {
  upload(d, v.empty() ? std::vector<T>(1) : v, st);
}
}  // namespace
}  // namespace tsem

using namespace tsem;

extern "C" {

int tsem_neumann(tsem_space* sp, const double* gq, int flag, double* b_dev) {
  API_BEGIN
  if (!sp || !gq || !b_dev) throw std::invalid_argument("null argument");
  Space& S = sp->h;
  if (S.mark[2].empty()) { g_err = "boundary marks not set"; return TSEM_ESTATE; }
  hipStream_t st = sp->st();
  bc_tables(sp);
  const int q2 = sp->nq2, nde = S.ndg[1], ndf = S.ndg[2];
  const int nloc = 3 + 3 * nde + ndf;
  auto faces = marked_geos(S, 2, flag);
  const int nf = (int)faces.size();
  if (nf == 0) return TSEM_OK;
  std::vector<double> area(nf);
  // target rows: (row, signed slot) in face order, grouped by row
  std::vector<std::pair<int64_t, int32_t>> ent;
  ent.reserve((size_t)nf * nloc);
  for (int jf = 0; jf < nf; ++jf) {
    const int f = faces[jf];
    const double* p0 = &S.xyz[3 * (size_t)S.face_node[f][0]];
    const double* p1 = &S.xyz[3 * (size_t)S.face_node[f][1]];
    const double* p2 = &S.xyz[3 * (size_t)S.face_node[f][2]];
    double u[3], v[3];
    for (int c = 0; c < 3; ++c) { u[c] = p1[c] - p0[c]; v[c] = p2[c] - p0[c]; }
    const double cx = u[1] * v[2] - u[2] * v[1], cy = u[0] * v[2] - u[2] * v[0], cz = u[0] * v[1] - u[1] * v[0];
    area[jf] = 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz);
    const int32_t base = jf * nloc;
    for (int j = 0; j < 3; ++j) ent.push_back({S.geo_dof(0, S.face_node[f][j]), base + j});
    for (int ie = 0; ie < 3; ++ie) {
      const bool rev = (S.face_rev[f] >> ie) & 1;
      const int64_t e0 = S.geo_dof(1, S.face_edges[f][ie]);
      for (int m = 0; m < nde; ++m) {
        int32_t sl = base + 3 + ie * nde + m;
        if (rev && (m % 2 == 1)) sl |= (int32_t)0x80000000;
        ent.push_back({e0 + m, sl});
      }
    }
    const int64_t fs = S.geo_dof(2, f);
    for (int dd = 0; dd < ndf; ++dd) ent.push_back({fs + dd, base + 3 + 3 * nde + dd});
  }
  std::stable_sort(ent.begin(), ent.end(),
                   [](const std::pair<int64_t, int32_t>& a, const std::pair<int64_t, int32_t>& b) {
                     return a.first < b.first;
                   });
  std::vector<int32_t> rows, ptr(1, 0), slots;
  for (size_t k = 0; k < ent.size(); ++k) {
    if (k == 0 || ent[k].first != ent[k - 1].first) {
      rows.push_back((int32_t)ent[k].first);
      if (k > 0) ptr.push_back((int32_t)k);
    }
    slots.push_back(ent[k].second);
  }
  ptr.push_back((int32_t)ent.size());
  DevBuf d_gq, d_area, d_contrib, d_rows, d_ptr, d_slots;
  std::vector<double> g(gq, gq + (size_t)nf * q2);
  put(d_gq, g, st);
  put(d_area, area, st);
  put(d_rows, rows, st);
  put(d_ptr, ptr, st);
  put(d_slots, slots, st);
  d_contrib.alloc((size_t)nf * nloc * 8);
  const long ns = (long)nf * nloc;
  hipLaunchKernelGGL(k_neu_local, dim3(nblk(ns)), dim3(BLK), 0, st, nf, nloc, nde, ndf, q2, d_area.as<double>(),
                     d_gq.as<double>(), sp->bc_tab.as<double>(), sp->bc_offd.as<long>(), d_contrib.as<double>());
  hipLaunchKernelGGL(k_neu_add, dim3(nblk((long)rows.size())), dim3(BLK), 0, st, (int)rows.size(),
                     d_rows.as<int32_t>(), d_ptr.as<int32_t>(), d_slots.as<int32_t>(), d_contrib.as<double>(),
                     b_dev);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(st));
  API_END
}

int tsem_dirichlet(tsem_space* sp, const double* uv, const double* ue, const double* uf, int flag, double* b_dev,
                   double* bval) {
  API_BEGIN
  if (!sp || !uv || !ue || !uf || !b_dev) throw std::invalid_argument("null argument");
  Space& S = sp->h;
  if (S.mark[0].empty()) { g_err = "boundary marks not set"; return TSEM_ESTATE; }
  if (S.ndg[1] > 11) throw std::invalid_argument("Dirichlet data need order <= 12");
  hipStream_t st = sp->st();
  bc_tables(sp);
  const int q1 = sp->nq1, q2 = sp->nq2, nde = S.ndg[1], ndf = S.ndg[2];
  auto gv = marked_geos(S, 0, flag), ge = marked_geos(S, 1, flag), gf = marked_geos(S, 2, flag);
  const int nv = (int)gv.size(), ne = (int)ge.size(), nf = (int)gf.size();
  std::vector<int32_t> vrow(nv), erow(ne), evs(ne), eve(ne), frow(nf), fv(3 * (size_t)nf), fe(3 * (size_t)nf),
      frev(nf);
  std::vector<uint8_t> newmask(S.ndof, 0);
  for (int j = 0; j < nv; ++j) {
    vrow[j] = (int32_t)S.geo_dof(0, gv[j]);
    newmask[vrow[j]] = 1;
  }
  for (int j = 0; j < ne; ++j) {
    const int g = ge[j];
    erow[j] = (int32_t)S.geo_dof(1, g);
    evs[j] = (int32_t)S.geo_dof(0, S.edge_node[g][0]);
    eve[j] = (int32_t)S.geo_dof(0, S.edge_node[g][1]);
    for (int m = 0; m < nde; ++m) newmask[erow[j] + m] = 1;
  }
  for (int j = 0; j < nf; ++j) {
    const int g = gf[j];
    frow[j] = (int32_t)S.geo_dof(2, g);
    for (int v = 0; v < 3; ++v) fv[3 * j + v] = (int32_t)S.geo_dof(0, S.face_node[g][v]);
    for (int ie = 0; ie < 3; ++ie) fe[3 * j + ie] = (int32_t)S.geo_dof(1, S.face_edges[g][ie]);
    frev[j] = S.face_rev[g];
    for (int dd = 0; dd < ndf; ++dd) newmask[frow[j] + dd] = 1;
  }
  DevBuf d_uv, d_ue, d_uf, d_vrow, d_erow, d_evs, d_eve, d_frow, d_fv, d_fe, d_frev, d_free, vb, w;
  put(d_uv, std::vector<double>(uv, uv + nv), st);
  put(d_ue, std::vector<double>(ue, ue + (size_t)ne * q1), st);
  put(d_uf, std::vector<double>(uf, uf + (size_t)nf * q2), st);
  put(d_vrow, vrow, st); put(d_erow, erow, st); put(d_evs, evs, st); put(d_eve, eve, st);
  put(d_frow, frow, st); put(d_fv, fv, st); put(d_fe, fe, st); put(d_frev, frev, st);
  const int n = (int)S.ndof;
  vb.alloc((size_t)n * 8);
  w.alloc((size_t)n * 8);
  launch_zero<double>(n, vb.as<double>(), st);
  const double* dg = sp->diag_d.as<double>();
  const double* tab = sp->bc_tab.as<double>();
  const long* off = sp->bc_offd.as<long>();
  if (nv > 0)
    hipLaunchKernelGGL(k_dir_vertex, dim3(nblk(nv)), dim3(BLK), 0, st, nv, d_vrow.as<int32_t>(), d_uv.as<double>(), dg,
                       b_dev, vb.as<double>());
  if (ne > 0 && nde > 0)
    hipLaunchKernelGGL(k_dir_edge, dim3(nblk((long)ne * nde)), dim3(BLK), 0, st, ne, nde, q1, d_erow.as<int32_t>(),
                       d_evs.as<int32_t>(), d_eve.as<int32_t>(), d_ue.as<double>(), tab, off, dg, b_dev,
                       vb.as<double>());
  if (nf > 0 && ndf > 0)
    hipLaunchKernelGGL(k_dir_face, dim3(nblk((long)nf * ndf)), dim3(BLK), 0, st, nf, ndf, nde, q2,
                       d_frow.as<int32_t>(), d_fv.as<int32_t>(), d_fe.as<int32_t>(), d_frev.as<int32_t>(),
                       d_uf.as<double>(), tab, off, dg, b_dev, vb.as<double>());
  // free rows: b -= A ext(val), with the elimination imposed so far on the
  // other DOFs (the DOFs imposed here are unmasked for this product, so data
  // imposed again on DOFs eliminated before still reach the free rows)
  std::vector<uint8_t> prev(sp->cmask), freerow(S.ndof);
  for (int64_t i = 0; i < S.ndof; ++i) {
    freerow[i] = !newmask[i] && !prev[i];
    if (newmask[i]) sp->cmask[i] = 0;
  }
  upload_mask(sp, st);
  put(d_free, freerow, st);
  OpCoef op;
  op.stiff = 1.0; op.mass = 0.0; op.constrained = 1;
  launch_apply<double>(sp->d, op, vb.as<double>(), w.as<double>(), sp->yl64.as<double>(), nullptr, st);
  hipLaunchKernelGGL(k_dir_fix, dim3(nblk(n)), dim3(BLK), 0, st, n, d_free.as<uint8_t>(), w.as<double>(), b_dev);
  bool grown = false;  // DOFs eliminated now that were free before
  for (int64_t i = 0; i < S.ndof; ++i) {
    grown |= newmask[i] && !prev[i];
    sp->cmask[i] = prev[i] | newmask[i];
  }
  upload_mask(sp, st);
  HIPCHK(hipGetLastError());
  if (bval) {
    HIPCHK(hipMemcpyAsync(bval, vb.p, (size_t)n * 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (int64_t i = 0; i < S.ndof; ++i)
      if (!newmask[i]) bval[i] = NAN;
  }
  HIPCHK(hipStreamSynchronize(st));
  // the p-multigrid levels and the captured solver graphs depend on the
  // elimination mask, not on the boundary values: with the mask as it was
  // (new data on parts eliminated before, e.g. each Hartree solve of an SCF
  // run) they stay valid
  if (grown) sp->drop_pmg();
  API_END
}

}  // extern "C"
