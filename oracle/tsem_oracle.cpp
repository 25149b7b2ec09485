// TEST INFRASTRUCTURE ONLY: CPU oracle for the TSEM operator and solver path.
// Only tests/, the smoke check and bench.py's cpu_baseline leg load the shared
// library built from this file; the product never links it.
//
// A restatement of TSEMLib's setup and solver, following:
//   src/BuildSpace.cpp:435-655    build_geometry_info (first-visit orientation)
//   src/BuildSpace.cpp:871-1101   build_global_transform (DOF layout, gather maps)
//   src/BuildSpace.cpp:1703-1771  read_interp_info (face table)
//   src/BuildSpace.cpp:60-147     2-D/1-D basis values and interpolation bases
//   src/Discretize.cpp:521-665    mass matrix assembly (coef 6V)
//   src/Discretize.cpp:1296-1395  stiffness assembly (coef 2/(3V), dihedral factors)
//   src/Discretize.cpp:1396-1418  calc_rhs from values at quadrature points
//   src/Discretize.cpp:1465-1551  impose_zero_boundary_condition
//   src/Discretize.cpp:1680-1780  Dirichlet modal interpolation + elimination
//   src/Discretize.cpp:2010-2063  Neumann face load
//   src/Miscellaneous.cpp:557-587 calc_sum_dof
//   src/TSEMSolver.cpp:35-136, 206-245, 344-382, 572-655  init_PMG and levels
//   src/TSEMSolver.cpp:923-1184   Gauss-Seidel sweeps, ite_PMG V-cycle
//   src/TSEMSolver.cpp:1775-1930  solve_PCG
// Reference element matrices are formed by exact collapsed Gauss-Legendre
// quadrature of the restated basis instead of the closed forms of
// src/Discretize.cpp:120-519, 844-1294 (same polynomials; tests pin them).
//
// The mesh topology and the P2 "fem dof" order that AFEPack supplies to the
// reference are fixed here by a canonical rule (see canonical_topology).
//
// Pins (tests/test_oracle.py): multi-index numbering against a table dumped
// from the reference's own Correspondence header; Jacobi modes against
// scipy; Gauss-Legendre against numpy; gradients against finite
// differences; reference matrices by independent quadrature, symmetry,
// null space and hierarchy; gathered fields continuous across faces; cubic
// solutions exact; PCG and the p-MG V-cycle against a direct sparse solve.
// Parity unpinned: AFEPack's absolute geometry numbering (replaced by the
// canonical rule), the closed-form matrix generators' rounding, deal.II's
// SpMV summation order, and the Gauss-Seidel V-cycle iterate by iterate
// (pinned only through the solution it converges to).
#include "oracle_basis.hpp"
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <numeric>
#include <sstream>
#include <string>
#include <tuple>

namespace oracle {

struct FaceTable {
  int mmax = 0, nface = 0;
  // [var][dof][edge k] -> (edge dof idx, w); [var][dof] -> (face dof idx, w)
  std::vector<std::vector<std::array<std::vector<std::pair<int, double>>, 3>>> edge;
  std::vector<std::vector<std::vector<std::pair<int, double>>>> face;
  bool load(const char* path) {
    std::ifstream in(path);
    if (!in) return false;
    std::string tag1, tag2;
    in >> tag1 >> mmax >> tag2 >> nface;
    edge.assign(6, std::vector<std::array<std::vector<std::pair<int, double>>, 3>>(nface));
    face.assign(6, std::vector<std::vector<std::pair<int, double>>>(nface));
    for (int r = 0; r < 6 * nface; ++r) {
      int var, dof;
      in >> var >> dof;
      for (int k = 0; k < 3; ++k) {
        int cnt; in >> cnt;
        for (int t = 0; t < cnt; ++t) { int i; double w; in >> i >> w; edge[var][dof][k].push_back({i, w}); }
      }
      int cnt; in >> cnt;
      for (int t = 0; t < cnt; ++t) { int i; double w; in >> i >> w; face[var][dof].push_back({i, w}); }
    }
    return (bool)in;
  }
};

// CSR with the deal.II convention: diagonal first in each row, then the other
// columns in increasing order.
struct CSR {
  int n = 0;
  std::vector<long> rs;
  std::vector<int> col;
  std::vector<double> v;
  double diag(int r) const { return v[rs[r]]; }
  long find(int r, int c) const {
    if (c == r) return rs[r];
    auto b = col.begin() + rs[r] + 1, e = col.begin() + rs[r + 1];
    auto it = std::lower_bound(b, e, c);
    if (it == e || *it != c) return -1;
    return it - col.begin();
  }
  void vmult(const double* x, double* y, int nthreads = 1) const {
#pragma omp parallel for schedule(static) num_threads(nthreads) if (nthreads > 1)
    for (int r = 0; r < n; ++r) {
      double s = 0;
      for (long k = rs[r]; k < rs[r + 1]; ++k) s += v[k] * x[col[k]];
      y[r] = s;
    }
  }
};

struct Space {
  // inputs
  int M = 0, nv = 0, nt = 0;
  std::vector<double> xyz;
  std::vector<int> tet;
  // canonical topology
  int ne = 0, nf = 0;
  std::vector<int> edge_vtx, face_vtx, face_edge, tet_face;
  std::vector<int> femdof_weight;  // per vertex then per edge
  // polynomial data
  Corr corr;
  LocalTransform lt;
  FaceTable ft;
  int n_index = 0, ndg[4] = {0, 0, 0, 0}, ngeo[4] = {0, 0, 0, 0}, n_geo_total = 0, ndof = 0;
  // geometry info
  std::vector<std::array<int, 4>> ig_v, ig_f;
  std::vector<std::array<int, 6>> ig_e;
  std::vector<std::array<int, 2>> num_node_e;
  std::vector<std::array<int, 3>> num_node_f, num_edge_f;
  std::vector<std::array<bool, 3>> sameorder;
  std::vector<std::array<int, 6>> tp_edge;
  std::vector<std::array<int, 4>> tp_face, tp_face_inv;
  std::vector<double> vol;
  // DOF layout
  std::vector<int> location_geometry, location_actualdof, geometry_dimension;
  // gather maps: per element, per local index: list of (global, w)
  std::vector<int> g2l_off;  // (nt*n_index + 1)
  std::vector<int> g2l_ind;
  std::vector<double> g2l_val;
  // reference element matrices (dense K x K, row-major)
  std::vector<double> Mref;
  std::vector<double> Sref[6];
  // RHS/BC quadrature and bases
  Quad quad;
  std::vector<double> bv3;        // [q3][n_index] actual basis at 3D points
  std::vector<double> bv2;        // [q2][ndg2] face basis (2D)
  std::vector<double> interp1;    // [q1][M-1]   J^{1,1}_{l-2}
  std::vector<double> interp2;    // [q2][ndg2]
  std::vector<double> addv;       // [q2][M-1][2]
  // boundary marks per geometry (vertex, edge, face)
  std::vector<int> mark[3];
  int geo_index(int dim, int i) const {
    int s = i;
    for (int d = 0; d < dim; ++d) s += ngeo[d];
    return s;
  }
  int dof_start(int dim, int i) const { return location_actualdof[location_geometry[geo_index(dim, i)]]; }
};

// --- canonical topology: faces opposite each vertex, edges E01..E23, numbered
// by first appearance scanning tets in order; P2 fem dofs numbered scanning
// tets in order, the 4 vertices then the 6 edges of each tet.
static const int FACE_PT[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
static const int EDGE_PT[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
static const int FACE_EDGE_LOCAL[4][3] = {{5, 4, 3}, {5, 2, 1}, {4, 2, 0}, {3, 1, 0}};

static void canonical_topology(Space& s) {
  std::map<std::array<int, 3>, int> fmap;
  std::map<std::array<int, 2>, int> emap;
  s.tet_face.assign(s.nt * 4, -1);
  auto edge_id = [&](int a, int b) {
    std::array<int, 2> key{std::min(a, b), std::max(a, b)};
    auto it = emap.find(key);
    if (it != emap.end()) return it->second;
    int id = (int)emap.size();
    emap[key] = id;
    s.edge_vtx.push_back(a); s.edge_vtx.push_back(b);
    return id;
  };
  for (int t = 0; t < s.nt; ++t) {
    const int* v = &s.tet[t * 4];
    for (int k = 0; k < 6; ++k) edge_id(v[EDGE_PT[k][0]], v[EDGE_PT[k][1]]);
  }
  s.ne = (int)emap.size();
  for (int t = 0; t < s.nt; ++t) {
    const int* v = &s.tet[t * 4];
    for (int k = 0; k < 4; ++k) {
      int a = v[FACE_PT[k][0]], b = v[FACE_PT[k][1]], c = v[FACE_PT[k][2]];
      std::array<int, 3> key{a, b, c};
      std::sort(key.begin(), key.end());
      auto it = fmap.find(key);
      int id;
      if (it == fmap.end()) {
        id = (int)fmap.size();
        fmap[key] = id;
        s.face_vtx.insert(s.face_vtx.end(), {a, b, c});
        // edge opposite face vertex j
        s.face_edge.push_back(emap.at({std::min(b, c), std::max(b, c)}));
        s.face_edge.push_back(emap.at({std::min(a, c), std::max(a, c)}));
        s.face_edge.push_back(emap.at({std::min(a, b), std::max(a, b)}));
      } else id = it->second;
      s.tet_face[t * 4 + k] = id;
    }
  }
  s.nf = (int)fmap.size();
}

static void canonical_femdof(Space& s) {
  std::map<std::array<int, 2>, int> emap;
  for (int e = 0; e < s.ne; ++e) {
    int a = s.edge_vtx[2 * e], b = s.edge_vtx[2 * e + 1];
    emap[{std::min(a, b), std::max(a, b)}] = e;
  }
  // one counter shared by vertices and edges, as a P2 DOF numbering built
  // element by element assigns them
  s.femdof_weight.assign(s.nv + s.ne, -1);
  int cnt = 0;
  for (int t = 0; t < s.nt; ++t) {
    const int* v = &s.tet[t * 4];
    for (int k = 0; k < 4; ++k)
      if (s.femdof_weight[v[k]] < 0) s.femdof_weight[v[k]] = cnt++;
    for (int k = 0; k < 6; ++k) {
      int a = v[EDGE_PT[k][0]], b = v[EDGE_PT[k][1]];
      int e = emap.at({std::min(a, b), std::max(a, b)});
      if (s.femdof_weight[s.nv + e] < 0) s.femdof_weight[s.nv + e] = cnt++;
    }
  }
}

// --- build_geometry_info (src/BuildSpace.cpp:435-655) ------------------------
static void geometry_info(Space& s) {
  s.ngeo[0] = s.nv; s.ngeo[1] = s.ne; s.ngeo[2] = s.nf; s.ngeo[3] = s.nt;
  s.n_geo_total = s.nv + s.ne + s.nf + s.nt;
  // C(M-1, d) DOFs per d-dimensional geometry.  The reference's binomial
  // helper returns 1 when M-1 < d (src/Miscellaneous.cpp:164-172), which gives
  // phantom face/interior DOFs for M <= 3; the count here is the true one.
  for (int d = 0; d <= 3; ++d) s.ndg[d] = (s.M - 1 >= d) ? Corr::binom(s.M - 1, d) : 0;
  s.ndof = 0;
  for (int d = 0; d <= 3; ++d) s.ndof += s.ngeo[d] * s.ndg[d];
  s.ig_v.resize(s.nt); s.ig_e.resize(s.nt); s.ig_f.resize(s.nt);
  s.num_node_e.assign(s.ne, {-1, -1});
  s.num_node_f.assign(s.nf, {-1, -1, -1});
  s.num_edge_f.assign(s.nf, {-1, -1, -1});
  s.sameorder.assign(s.nf, {true, true, true});
  s.tp_edge.assign(s.nt, {-1, -1, -1, -1, -1, -1});
  s.tp_face.assign(s.nt, {-1, -1, -1, -1});
  s.tp_face_inv.assign(s.nt, {-1, -1, -1, -1});
  std::vector<char> fe(s.ne, 0), ff(s.nf, 0);
  static const int SP[6] = {0, 0, 0, 1, 1, 2}, EP[6] = {1, 2, 3, 2, 3, 3};
  for (int e = 0; e < s.nt; ++e) {
    for (int k = 0; k < 4; ++k) s.ig_v[e][k] = s.tet[e * 4 + k];
    auto& V = s.ig_v[e];
    auto& Eg = s.ig_e[e];
    // edges via the faces opposite P3, P2, P1
    int f = s.tet_face[e * 4 + 3];
    for (int k = 0; k < 3; ++k) {
      int fv = s.face_vtx[3 * f + k], fe_ = s.face_edge[3 * f + k];
      if (fv == V[0]) Eg[3] = fe_;
      if (fv == V[1]) Eg[1] = fe_;
      if (fv == V[2]) Eg[0] = fe_;
    }
    f = s.tet_face[e * 4 + 2];
    for (int k = 0; k < 3; ++k) {
      int fv = s.face_vtx[3 * f + k], fe_ = s.face_edge[3 * f + k];
      if (fv == V[0]) Eg[4] = fe_;
      if (fv == V[1]) Eg[2] = fe_;
    }
    f = s.tet_face[e * 4 + 1];
    for (int k = 0; k < 3; ++k)
      if (s.face_vtx[3 * f + k] == V[0]) Eg[5] = s.face_edge[3 * f + k];
    for (int k = 0; k < 4; ++k) s.ig_f[e][k] = s.tet_face[e * 4 + k];
    for (int k = 0; k < 6; ++k) {
      int ie = Eg[k];
      if (!fe[ie]) {
        s.num_node_e[ie] = {V[SP[k]], V[EP[k]]};
        s.tp_edge[e][k] = 0;
        fe[ie] = 1;
      } else {
        s.tp_edge[e][k] = (s.num_node_e[ie][0] == V[SP[k]]) ? 0 : 1;
      }
    }
    for (int k = 0; k < 4; ++k) {
      int iff = s.ig_f[e][k];
      if (!ff[iff]) {
        for (int j = 0; j < 3; ++j) {
          s.num_node_f[iff][j] = V[FACE_PT[k][j]];
          s.num_edge_f[iff][j] = Eg[FACE_EDGE_LOCAL[k][j]];
        }
        if (s.num_node_e[Eg[FACE_EDGE_LOCAL[k][0]]][0] != V[FACE_PT[k][1]]) s.sameorder[iff][0] = false;
        if (s.num_node_e[Eg[FACE_EDGE_LOCAL[k][1]]][0] != V[FACE_PT[k][0]]) s.sameorder[iff][1] = false;
        if (s.num_node_e[Eg[FACE_EDGE_LOCAL[k][2]]][0] != V[FACE_PT[k][0]]) s.sameorder[iff][2] = false;
        s.tp_face[e][k] = 0; s.tp_face_inv[e][k] = 0;
        ff[iff] = 1;
      } else {
        int ord[3] = {-1, -1, -1};
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b)
            if (V[FACE_PT[k][a]] == s.num_node_f[iff][b]) ord[a] = b;
        int flag = ord[0] * 100 + ord[1] * 10 + ord[2];
        int tp = -1, tpi = -1;
        switch (flag) {
          case 12: tp = 0; tpi = 0; break;
          case 21: tp = 1; tpi = 1; break;
          case 102: tp = 2; tpi = 2; break;
          case 120: tp = 3; tpi = 4; break;
          case 201: tp = 4; tpi = 3; break;
          case 210: tp = 5; tpi = 5; break;
        }
        s.tp_face[e][k] = tp; s.tp_face_inv[e][k] = tpi;
      }
    }
  }
  s.vol.resize(s.nt);
  for (int e = 0; e < s.nt; ++e) {
    const double* p0 = &s.xyz[3 * s.ig_v[e][0]];
    const double* p1 = &s.xyz[3 * s.ig_v[e][1]];
    const double* p2 = &s.xyz[3 * s.ig_v[e][2]];
    const double* p3 = &s.xyz[3 * s.ig_v[e][3]];
    double a[3], b[3], c[3];
    for (int i = 0; i < 3; ++i) { a[i] = p1[i] - p0[i]; b[i] = p2[i] - p0[i]; c[i] = p3[i] - p0[i]; }
    double det = a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0]) + a[2] * (b[0] * c[1] - b[1] * c[0]);
    s.vol[e] = std::fabs(det) / 6.0;
  }
}

// --- build_global_transform (src/BuildSpace.cpp:871-1101) --------------------
static void global_transform(Space& s) {
  int n = s.n_geo_total;
  std::vector<double> w(n);
  s.geometry_dimension.resize(n);
  std::vector<int> gorder(n);
  int st = 0;
  for (int d = 0; d <= 3; ++d) {
    for (int i = 0; i < s.ngeo[d]; ++i) { s.geometry_dimension[st + i] = d; gorder[st + i] = i; }
    st += s.ngeo[d];
  }
  for (int i = 0; i < s.nv; ++i) w[i] = s.femdof_weight[i];
  for (int i = 0; i < s.ne; ++i) w[s.nv + i] = s.femdof_weight[s.nv + i];
  for (int i = 0; i < s.nf; ++i) {
    double t = 0;
    for (int k = 0; k < 3; ++k) t += w[s.face_vtx[3 * i + k]];
    w[s.nv + s.ne + i] = t / 3;
  }
  for (int i = 0; i < s.nt; ++i) {
    double t = 0;
    for (int k = 0; k < 4; ++k) t += w[s.tet[4 * i + k]];
    w[s.nv + s.ne + s.nf + i] = t / 4;
  }
  // stable sort by weight (the reference's insertion sort is stable)
  std::vector<int> perm(n);
  std::iota(perm.begin(), perm.end(), 0);
  std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return w[a] < w[b]; });
  std::vector<int> dim_sorted(n);
  s.location_geometry.assign(n, -1);
  for (int i = 0; i < n; ++i) { s.location_geometry[perm[i]] = i; dim_sorted[i] = s.geometry_dimension[perm[i]]; }
  s.location_actualdof.assign(n, 0);
  for (int i = 1; i < n; ++i) s.location_actualdof[i] = s.location_actualdof[i - 1] + s.ndg[dim_sorted[i - 1]];

  const Corr& c = s.corr;
  int M = s.M, K = s.n_index, nde = M - 1, ndf = s.ndg[2];
  std::vector<int> edgedof(6 * nde), facedof(4 * std::max(ndf, 1));
  for (int l = 2; l <= M; ++l) {
    edgedof[nde * 0 + l - 2] = c.i2n(l, 0, 0);
    edgedof[nde * 1 + l - 2] = c.i2n(0, l, 0);
    edgedof[nde * 2 + l - 2] = c.i2n(0, 0, l);
    edgedof[nde * 4 + l - 2] = c.i2n(1, 0, l - 1);
    edgedof[nde * 3 + l - 2] = c.i2n(1, l - 1, 0);
    edgedof[nde * 5 + l - 2] = c.i2n(0, 1, l - 1);
  }
  for (int l1 = 2; l1 <= M; ++l1)
    for (int l2 = 1; l2 <= M - l1; ++l2) {
      int id = (l1 + l2 - 3) * (l1 + l2 - 2) / 2 + l2 - 1;
      facedof[ndf * 0 + id] = c.i2n(1, l1 - 1, l2);
      facedof[ndf * 1 + id] = c.i2n(0, l1, l2);
      facedof[ndf * 2 + id] = c.i2n(l1, 0, l2);
      facedof[ndf * 3 + id] = c.i2n(l1, l2, 0);
    }
  std::vector<std::vector<std::pair<int, double>>> loc(K);
  s.g2l_off.assign((size_t)s.nt * K + 1, 0);
  s.g2l_ind.clear(); s.g2l_val.clear();
  for (int e = 0; e < s.nt; ++e) {
    for (auto& v : loc) v.clear();
    for (int k = 0; k < 4; ++k) loc[k].push_back({s.dof_start(0, s.ig_v[e][k]), 1.0});
    for (int k = 0; k < 6; ++k) {
      int st0 = s.dof_start(1, s.ig_e[e][k]);
      for (int l = 2; l <= M; ++l) {
        double v = 1.0;
        if (s.tp_edge[e][k] == 1 && l % 2 == 1) v = -1.0;
        loc[edgedof[nde * k + l - 2]].push_back({st0 + l - 2, v});
      }
    }
    for (int k = 0; k < 4; ++k) {
      int f = s.ig_f[e][k];
      int fst = s.dof_start(2, f);
      int var = s.tp_face_inv[e][k];
      for (int d = 0; d < ndf; ++d) {
        auto& L = loc[facedof[ndf * k + d]];
        for (int ie = 0; ie < 3; ++ie) {
          int est = s.dof_start(1, s.num_edge_f[f][ie]);
          for (auto& pr : s.ft.edge[var][d][ie]) {
            double v = pr.second;
            if (!s.sameorder[f][ie] && pr.first % 2 == 1) v *= -1;
            L.push_back({est + pr.first, v});
          }
        }
        for (auto& pr : s.ft.face[var][d]) L.push_back({fst + pr.first, pr.second});
      }
    }
    int ist = s.dof_start(3, e);
    for (int d = 0; d < s.ndg[3]; ++d) {
      auto a = c.n2i[d];
      loc[c.i2n(a[0] + 2, a[1] + 1, a[2] + 1)].push_back({ist + d, 1.0});
    }
    for (int i = 0; i < K; ++i) {
      for (auto& pr : loc[i]) { s.g2l_ind.push_back(pr.first); s.g2l_val.push_back(pr.second); }
      s.g2l_off[(size_t)e * K + i + 1] = (int)s.g2l_ind.size();
    }
  }
}

// --- reference matrices by quadrature -----------------------------------------
static void reference_matrices(Space& s) {
  int K = s.n_index;
  Quad q;
  q.build(s.M + 3);
  s.Mref.assign((size_t)K * K, 0.0);
  for (int k = 0; k < 6; ++k) s.Sref[k].assign((size_t)K * K, 0.0);
  std::vector<D3> phi;
  std::vector<double> val(K), dk[6];
  for (int k = 0; k < 6; ++k) dk[k].resize(K);
  for (size_t p = 0; p < q.p3.size(); ++p) {
    D3 x(q.p3[p][0]), y(q.p3[p][1]), z(q.p3[p][2]);
    x.d[0] = 1; y.d[1] = 1; z.d[2] = 1;
    actual_basis<D3>(s.corr, s.lt, x, y, z, phi);
    double w = q.w3[p] / 6.0;
    for (int i = 0; i < K; ++i) {
      val[i] = phi[i].v;
      const double* g = phi[i].d;
      dk[0][i] = g[0];          // d1        (2,0,0)
      dk[1][i] = g[1] - g[0];   // d2 - d1   (1,1,0)
      dk[2][i] = g[0] - g[2];   // d1 - d3   (1,0,1)
      dk[3][i] = g[1];          // d2        (0,2,0)
      dk[4][i] = g[2] - g[1];   // d3 - d2   (0,1,1)
      dk[5][i] = g[2];          // d3        (0,0,2)
    }
    for (int i = 0; i < K; ++i)
      for (int j = 0; j < K; ++j) {
        s.Mref[(size_t)i * K + j] += w * val[i] * val[j];
        for (int k = 0; k < 6; ++k) s.Sref[k][(size_t)i * K + j] += w * dk[k][i] * dk[k][j];
      }
  }
}

// per-element stiffness pair weights (src/Discretize.cpp:1355-1370): returns
// 6 coefficients indexed like Sref (0..5), zero when the dihedral factor is
// below 1e-12 in magnitude.
static void stiff_weights(const Space& s, int e, double* wk) {
  for (int k = 0; k < 6; ++k) wk[k] = 0;
  double coef = 2.0 / (3 * s.vol[e]);
  for (int i1 = 0; i1 <= 2; ++i1)
    for (int i2 = i1 + 1; i2 <= 3; ++i2) {
      int ip1 = (i1 == 0) ? (i2 % 2 + 1) : 0;
      int ip2 = 6 - i1 - i2 - ip1;
      const double* p1 = &s.xyz[3 * s.ig_v[e][i1]];
      const double* p2 = &s.xyz[3 * s.ig_v[e][i2]];
      const double* ps = &s.xyz[3 * s.ig_v[e][ip1]];
      const double* pe = &s.xyz[3 * s.ig_v[e][ip2]];
      double a[3], b[3], c[3];
      for (int i = 0; i < 3; ++i) { a[i] = pe[i] - ps[i]; b[i] = p1[i] - ps[i]; c[i] = p2[i] - ps[i]; }
      double u[3] = {a[1] * b[2] - a[2] * b[1], -a[0] * b[2] + a[2] * b[0], a[0] * b[1] - a[1] * b[0]};
      double v[3] = {a[1] * c[2] - a[2] * c[1], -a[0] * c[2] + a[2] * c[0], a[0] * c[1] - a[1] * c[0]};
      double dih = 0.25 * (u[0] * v[0] + u[1] * v[1] + u[2] * v[2]);
      if (std::fabs(dih) < 1e-12) continue;
      int ind;
      if (i1 == 0) ind = s.corr.i2n(i2 == 1 ? 2 : 0, i2 == 2 ? 2 : 0, i2 == 3 ? 2 : 0) - 4;
      else {
        int m[3] = {0, 0, 0};
        m[i1 - 1] += 1; m[i2 - 1] += 1;
        ind = s.corr.i2n(m[0], m[1], m[2]) - 4;
      }
      wk[ind] += coef * dih;
    }
}

// dense element matrix K_e = sum_k w_k S_k + mass_coef * 6V * Mref
static void element_matrix(const Space& s, int e, double stiff_coef, double mass_coef, std::vector<double>& Ke) {
  int K = s.n_index;
  Ke.assign((size_t)K * K, 0.0);
  double wk[6];
  stiff_weights(s, e, wk);
  for (int k = 0; k < 6; ++k) {
    if (wk[k] == 0 || stiff_coef == 0) continue;
    for (size_t t = 0; t < (size_t)K * K; ++t) Ke[t] += stiff_coef * wk[k] * s.Sref[k][t];
  }
  if (mass_coef != 0)
    for (size_t t = 0; t < (size_t)K * K; ++t) Ke[t] += mass_coef * 6 * s.vol[e] * s.Mref[t];
}

// global CSR assembly sum_e G_e^T K_e G_e (src/Discretize.cpp:1296-1395)
static CSR assemble(const Space& s, double stiff_coef, double mass_coef) {
  int K = s.n_index;
  std::vector<std::vector<int>> cols(s.ndof);
  for (int e = 0; e < s.nt; ++e) {
    std::vector<int> g;
    for (int i = 0; i < K; ++i)
      for (int t = s.g2l_off[(size_t)e * K + i]; t < s.g2l_off[(size_t)e * K + i + 1]; ++t) g.push_back(s.g2l_ind[t]);
    std::sort(g.begin(), g.end());
    g.erase(std::unique(g.begin(), g.end()), g.end());
    for (int r : g) cols[r].insert(cols[r].end(), g.begin(), g.end());
  }
  CSR A;
  A.n = s.ndof;
  A.rs.assign(s.ndof + 1, 0);
  for (int r = 0; r < s.ndof; ++r) {
    auto& c = cols[r];
    std::sort(c.begin(), c.end());
    c.erase(std::unique(c.begin(), c.end()), c.end());
    A.rs[r + 1] = A.rs[r] + (long)c.size();
  }
  A.col.resize(A.rs[s.ndof]);
  A.v.assign(A.rs[s.ndof], 0.0);
  for (int r = 0; r < s.ndof; ++r) {
    long p = A.rs[r];
    A.col[p++] = r;
    for (int c : cols[r]) if (c != r) A.col[p++] = c;
    std::vector<int>().swap(cols[r]);
  }
  std::vector<double> Ke;
  for (int e = 0; e < s.nt; ++e) {
    element_matrix(s, e, stiff_coef, mass_coef, Ke);
    for (int i = 0; i < K; ++i)
      for (int j = 0; j < K; ++j) {
        double kv = Ke[(size_t)i * K + j];
        if (kv == 0) continue;
        for (int a = s.g2l_off[(size_t)e * K + i]; a < s.g2l_off[(size_t)e * K + i + 1]; ++a)
          for (int b = s.g2l_off[(size_t)e * K + j]; b < s.g2l_off[(size_t)e * K + j + 1]; ++b) {
            long pos = A.find(s.g2l_ind[a], s.g2l_ind[b]);
            A.v[pos] += kv * s.g2l_val[a] * s.g2l_val[b];
          }
      }
  }
  return A;
}

// --- 2-D/1-D bases for BC (src/BuildSpace.cpp:60-147) -------------------------
static void bc_bases(Space& s, int qorder) {
  int M = s.M, nde = M - 1, ndf = s.ndg[2];
  s.quad.build(qorder);
  const Quad& q = s.quad;
  int q1 = (int)q.t1.size(), q2 = (int)q.p2.size();
  s.bv2.assign((size_t)q2 * std::max(ndf, 1), 0.0);
  s.interp1.assign((size_t)q1 * nde, 0.0);
  s.interp2.assign((size_t)q2 * std::max(ndf, 1), 0.0);
  s.addv.assign((size_t)q2 * nde * 2, 0.0);
  std::vector<double> Jxi(M + 2), Jeta(M + 2);
  for (int p = 0; p < q2; ++p) {
    double x = q.p2[p][0], y = q.p2[p][1];
    double xi = 2 * x / (1 - y) - 1, eta = 2 * y - 1;
    jac_rec<double>(-1, -1, M, xi, Jxi.data());
    for (int l1 = 2; l1 <= M; ++l1) {
      jac_rec<double>(2 * l1 - 1, -1, M, eta, Jeta.data());
      for (int l2 = 1; l2 <= M - l1; ++l2) {
        int id = (l1 + l2 - 2) * (l1 + l2 - 3) / 2 + l2 - 1;
        s.bv2[(size_t)p * ndf + id] = 2 * std::pow(1 - y, l1) * Jxi[l1] * Jeta[l2];
      }
    }
  }
  for (int p = 0; p < q1; ++p) {
    double xp = 2 * q.t1[p] - 1;
    if (M >= 2) {
      std::vector<double> J(M + 1);
      jac_rec<double>(1, 1, std::max(M - 2, 0), xp, J.data());
      for (int l = 0; l <= M - 2; ++l) s.interp1[(size_t)p * nde + l] = J[l];
    }
  }
  for (int p = 0; p < q2; ++p) {
    double xp = q.p2[p][0], yp = q.p2[p][1];
    double xi = 2 * xp / (1 - yp) - 1, eta = 2 * yp - 1;
    std::vector<double> Jx(M + 1), Je(M + 1);
    jac_rec<double>(1, 1, std::max(M - 2, 0), xi, Jx.data());
    for (int l1 = 2; l1 <= M; ++l1) {
      jac_rec<double>(2 * l1 - 1, 1, std::max(M - l1 - 1, 0), eta, Je.data());
      for (int l2 = 1; l2 <= M - l1; ++l2) {
        int id = (l1 + l2 - 2) * (l1 + l2 - 3) / 2 + l2 - 1;
        s.interp2[(size_t)p * ndf + id] = std::pow(1 - eta, l1 - 2) * Jx[l1 - 2] * Je[l2 - 1];
      }
    }
    for (int l = 2; l <= M; ++l) s.addv[((size_t)p * nde + l - 2) * 2 + 0] = std::pow(1 - yp, l - 2) * Jx[l - 2];
    jac_rec<double>(1, 1, std::max(M - 2, 0), eta, Je.data());
    for (int l = 2; l <= M; ++l) s.addv[((size_t)p * nde + l - 2) * 2 + 1] = Je[l - 2];
  }
}

}  // namespace oracle

using namespace oracle;

// ============================================================================
// C API (ctypes)
// ============================================================================
extern "C" {

struct OrHandle {
  Space s;
  std::unique_ptr<CSR> A;   // last assembled (possibly BC-modified) operator
};

void* or_space_create(int M, int nv, const double* xyz, int nt, const int* tet, const char* face_table,
                      const int* femdof_weight, int qorder_bc) {
  auto* h = new OrHandle();
  Space& s = h->s;
  s.M = M; s.nv = nv; s.nt = nt;
  s.xyz.assign(xyz, xyz + 3 * nv);
  s.tet.assign(tet, tet + 4 * nt);
  s.corr.init(M);
  s.n_index = s.corr.n;
  s.lt.build(s.corr);
  if (!s.ft.load(face_table) || s.ft.mmax < M) { delete h; return nullptr; }
  canonical_topology(s);
  if (femdof_weight) s.femdof_weight.assign(femdof_weight, femdof_weight + nv + s.ne);
  else canonical_femdof(s);
  geometry_info(s);
  global_transform(s);
  reference_matrices(s);
  bc_bases(s, qorder_bc);
  return h;
}

void or_space_destroy(void* p) { delete (OrHandle*)p; }

// info: [n_dof, n_index, nv, ne, nf, nt, ndg0..3, nnz_g2l]
void or_info(void* p, long* out) {
  Space& s = ((OrHandle*)p)->s;
  out[0] = s.ndof; out[1] = s.n_index; out[2] = s.nv; out[3] = s.ne; out[4] = s.nf; out[5] = s.nt;
  for (int d = 0; d < 4; ++d) out[6 + d] = s.ndg[d];
  out[10] = (long)s.g2l_ind.size();
}

void or_topology(void* p, int* edge_vtx, int* face_vtx, int* face_edge, int* tet_face) {
  Space& s = ((OrHandle*)p)->s;
  std::copy(s.edge_vtx.begin(), s.edge_vtx.end(), edge_vtx);
  std::copy(s.face_vtx.begin(), s.face_vtx.end(), face_vtx);
  std::copy(s.face_edge.begin(), s.face_edge.end(), face_edge);
  std::copy(s.tet_face.begin(), s.tet_face.end(), tet_face);
}

void or_maps(void* p, int* off, int* ind, double* val) {
  Space& s = ((OrHandle*)p)->s;
  std::copy(s.g2l_off.begin(), s.g2l_off.end(), off);
  std::copy(s.g2l_ind.begin(), s.g2l_ind.end(), ind);
  std::copy(s.g2l_val.begin(), s.g2l_val.end(), val);
}

// per geometry (vertex, edge, face, tet) first global DOF; geometry dims in order
void or_dof_start(void* p, int* out) {
  Space& s = ((OrHandle*)p)->s;
  for (int g = 0; g < s.n_geo_total; ++g) out[g] = s.location_actualdof[s.location_geometry[g]];
}

// orientation data: tp_edge[nt*6], tp_face_inv[nt*4], sameorder[nf*3], num_node_f[nf*3], num_edge_f[nf*3], num_node_e[ne*2]
void or_orientation(void* p, int* tp_edge, int* tp_face_inv, int* sameorder, int* num_node_f, int* num_edge_f, int* num_node_e) {
  Space& s = ((OrHandle*)p)->s;
  for (int e = 0; e < s.nt; ++e) {
    for (int k = 0; k < 6; ++k) tp_edge[e * 6 + k] = s.tp_edge[e][k];
    for (int k = 0; k < 4; ++k) tp_face_inv[e * 4 + k] = s.tp_face_inv[e][k];
  }
  for (int f = 0; f < s.nf; ++f)
    for (int k = 0; k < 3; ++k) {
      sameorder[f * 3 + k] = s.sameorder[f][k] ? 1 : 0;
      num_node_f[f * 3 + k] = s.num_node_f[f][k];
      num_edge_f[f * 3 + k] = s.num_edge_f[f][k];
    }
  for (int e = 0; e < s.ne; ++e) { num_node_e[2 * e] = s.num_node_e[e][0]; num_node_e[2 * e + 1] = s.num_node_e[e][1]; }
}

void or_volumes(void* p, double* out) {
  Space& s = ((OrHandle*)p)->s;
  std::copy(s.vol.begin(), s.vol.end(), out);
}

void or_ref_matrices(void* p, double* Mref, double* S6) {
  Space& s = ((OrHandle*)p)->s;
  size_t KK = (size_t)s.n_index * s.n_index;
  std::copy(s.Mref.begin(), s.Mref.end(), Mref);
  for (int k = 0; k < 6; ++k) std::copy(s.Sref[k].begin(), s.Sref[k].end(), S6 + k * KK);
}

void or_stiff_weights(void* p, double* out) {
  Space& s = ((OrHandle*)p)->s;
  for (int e = 0; e < s.nt; ++e) stiff_weights(s, e, out + 6 * e);
}

// actual basis values and gradients at reference points (n points)
void or_basis_at(void* p, int n, const double* pts, double* val, double* grad) {
  Space& s = ((OrHandle*)p)->s;
  int K = s.n_index;
  std::vector<D3> phi;
  for (int i = 0; i < n; ++i) {
    D3 x(pts[3 * i]), y(pts[3 * i + 1]), z(pts[3 * i + 2]);
    x.d[0] = 1; y.d[1] = 1; z.d[2] = 1;
    actual_basis<D3>(s.corr, s.lt, x, y, z, phi);
    for (int k = 0; k < K; ++k) {
      val[(size_t)i * K + k] = phi[k].v;
      for (int d = 0; d < 3; ++d) grad[((size_t)i * K + k) * 3 + d] = phi[k].d[d];
    }
  }
}

void or_multiindex(void* p, int* out) {
  Space& s = ((OrHandle*)p)->s;
  for (int i = 0; i < s.n_index; ++i)
    for (int d = 0; d < 3; ++d) out[3 * i + d] = s.corr.n2i[i][d];
}

// matrix-free reference apply y = sum_e G^T K_e G x (no BC), for small meshes
void or_apply(void* p, double stiff_coef, double mass_coef, const double* x, double* y) {
  Space& s = ((OrHandle*)p)->s;
  int K = s.n_index;
  std::fill(y, y + s.ndof, 0.0);
  std::vector<double> Ke, xl(K), yl(K);
  for (int e = 0; e < s.nt; ++e) {
    element_matrix(s, e, stiff_coef, mass_coef, Ke);
    for (int i = 0; i < K; ++i) {
      double t = 0;
      for (int a = s.g2l_off[(size_t)e * K + i]; a < s.g2l_off[(size_t)e * K + i + 1]; ++a) t += s.g2l_val[a] * x[s.g2l_ind[a]];
      xl[i] = t;
    }
    for (int i = 0; i < K; ++i) {
      double t = 0;
      for (int j = 0; j < K; ++j) t += Ke[(size_t)i * K + j] * xl[j];
      yl[i] = t;
    }
    for (int i = 0; i < K; ++i)
      for (int a = s.g2l_off[(size_t)e * K + i]; a < s.g2l_off[(size_t)e * K + i + 1]; ++a) y[s.g2l_ind[a]] += s.g2l_val[a] * yl[i];
  }
}

// gather: local coefficients of every element, xl[nt*K]
void or_gather(void* p, const double* x, double* xl) {
  Space& s = ((OrHandle*)p)->s;
  int K = s.n_index;
  for (size_t t = 0; t < (size_t)s.nt * K; ++t) {
    double v = 0;
    for (int a = s.g2l_off[t]; a < s.g2l_off[t + 1]; ++a) v += x[s.g2l_ind[a]] * s.g2l_val[a];
    xl[t] = v;
  }
}

// assemble the global CSR (kept in the handle); returns nnz
long or_assemble(void* p, double stiff_coef, double mass_coef) {
  OrHandle* h = (OrHandle*)p;
  h->A.reset(new CSR(assemble(h->s, stiff_coef, mass_coef)));
  return (long)h->A->v.size();
}

void or_csr(void* p, long* rs, int* col, double* v) {
  OrHandle* h = (OrHandle*)p;
  std::copy(h->A->rs.begin(), h->A->rs.end(), rs);
  std::copy(h->A->col.begin(), h->A->col.end(), col);
  std::copy(h->A->v.begin(), h->A->v.end(), v);
}

void or_csr_vmult(void* p, const double* x, double* y) { ((OrHandle*)p)->A->vmult(x, y); }

// boundary marks per geometry (vertex, edge, face), as build_flag_bm(mesh)
void or_set_marks(void* p, const int* mv, const int* me, const int* mf) {
  Space& s = ((OrHandle*)p)->s;
  s.mark[0].assign(mv, mv + s.nv);
  s.mark[1].assign(me, me + s.ne);
  s.mark[2].assign(mf, mf + s.nf);
}

// number of BC quadrature points: q1, q2, q3(used for rhs)
void or_bc_quad(void* p, int* nq, double* t1, double* w1, double* p2, double* w2) {
  Space& s = ((OrHandle*)p)->s;
  const Quad& q = s.quad;
  nq[0] = (int)q.t1.size(); nq[1] = (int)q.p2.size(); nq[2] = (int)q.p3.size();
  if (t1) { std::copy(q.t1.begin(), q.t1.end(), t1); std::copy(q.w1.begin(), q.w1.end(), w1); }
  if (p2) {
    for (size_t i = 0; i < q.p2.size(); ++i) for (int d = 0; d < 3; ++d) p2[3 * i + d] = q.p2[i][d];
    std::copy(q.w2.begin(), q.w2.end(), w2);
  }
}

void or_quad3(void* p, double* pts, double* w) {
  Space& s = ((OrHandle*)p)->s;
  const Quad& q = s.quad;
  for (size_t i = 0; i < q.p3.size(); ++i) for (int d = 0; d < 3; ++d) pts[3 * i + d] = q.p3[i][d];
  std::copy(q.w3.begin(), q.w3.end(), w);
}

// calc_rhs(rhs, value_f) (src/Discretize.cpp:1396-1418); fq[nt][nq3]
void or_rhs(void* p, const double* fq, double* rhs) {
  Space& s = ((OrHandle*)p)->s;
  int K = s.n_index;
  const Quad& q = s.quad;
  size_t nq = q.p3.size();
  if (s.bv3.empty()) {
    s.bv3.assign(nq * K, 0.0);
    std::vector<double> phi;
    for (size_t i = 0; i < nq; ++i) {
      actual_basis<double>(s.corr, s.lt, q.p3[i][0], q.p3[i][1], q.p3[i][2], phi);
      for (int k = 0; k < K; ++k) s.bv3[i * K + k] = phi[k];
    }
  }
  std::fill(rhs, rhs + s.ndof, 0.0);
  for (int e = 0; e < s.nt; ++e)
    for (int i = 0; i < K; ++i) {
      double cnt = 0;
      for (size_t pq = 0; pq < nq; ++pq) cnt += q.w3[pq] * fq[(size_t)e * nq + pq] * s.bv3[pq * K + i];
      cnt *= s.vol[e];
      for (int a = s.g2l_off[(size_t)e * K + i]; a < s.g2l_off[(size_t)e * K + i + 1]; ++a) rhs[s.g2l_ind[a]] += cnt * s.g2l_val[a];
    }
}

// physical coordinates of 3D quadrature points of every element, out[nt][nq3][3]
void or_qp_coords(void* p, double* out) {
  Space& s = ((OrHandle*)p)->s;
  const Quad& q = s.quad;
  size_t nq = q.p3.size();
  for (int e = 0; e < s.nt; ++e) {
    const double* P[4];
    for (int k = 0; k < 4; ++k) P[k] = &s.xyz[3 * s.ig_v[e][k]];
    for (size_t i = 0; i < nq; ++i) {
      double x = q.p3[i][0], y = q.p3[i][1], z = q.p3[i][2];
      double l0 = 1 - x - y - z;
      for (int d = 0; d < 3; ++d)
        out[((size_t)e * nq + i) * 3 + d] = l0 * P[0][d] + x * P[1][d] + y * P[2][d] + z * P[3][d];
    }
  }
}

// Neumann load on faces with mark==flag (src/Discretize.cpp:2010-2063).
// g given at face quadrature points: gq[nf][nq2] (entries of other faces unused).
void or_neumann(void* p, const double* gq, int flag, double* rhs) {
  Space& s = ((OrHandle*)p)->s;
  const Quad& q = s.quad;
  int q2 = (int)q.p2.size(), nde = s.ndg[1], ndf = s.ndg[2];
  for (int f = 0; f < s.nf; ++f) {
    if (s.mark[2][f] != flag) continue;
    double cv[3] = {0, 0, 0};
    std::vector<std::vector<double>> ce(3, std::vector<double>(nde, 0.0));
    std::vector<double> cf(ndf, 0.0);
    const double* p0 = &s.xyz[3 * s.num_node_f[f][0]];
    const double* p1 = &s.xyz[3 * s.num_node_f[f][1]];
    const double* p2 = &s.xyz[3 * s.num_node_f[f][2]];
    double area = 0.5 * std::sqrt(std::pow((p1[1] - p0[1]) * (p2[2] - p0[2]) - (p1[2] - p0[2]) * (p2[1] - p0[1]), 2) +
                                  std::pow((p1[0] - p0[0]) * (p2[2] - p0[2]) - (p1[2] - p0[2]) * (p2[0] - p0[0]), 2) +
                                  std::pow((p1[0] - p0[0]) * (p2[1] - p0[1]) - (p1[1] - p0[1]) * (p2[0] - p0[0]), 2));
    for (int qp = 0; qp < q2; ++qp) {
      const auto& qb = q.p2[qp];
      double wgv = q.w2[qp] * gq[(size_t)f * q2 + qp] * area;
      for (int v = 0; v < 3; ++v) cv[v] += wgv * qb[(v + 2) % 3];
      double xp = qb[0], yp = qb[1], rp = qb[2];
      for (int d = 0; d < nde; ++d) {
        const double* ad = &s.addv[((size_t)qp * nde + d) * 2];
        ce[0][d] += wgv * -2 * xp * yp * ad[1];
        ce[1][d] += wgv * -2 * yp * rp * ad[1];
        ce[2][d] += wgv * -2 * xp * rp * ad[0];
      }
      for (int d = 0; d < ndf; ++d) cf[d] += wgv * s.bv2[(size_t)qp * ndf + d];
    }
    for (int v = 0; v < 3; ++v) rhs[s.dof_start(0, s.num_node_f[f][v])] += cv[v];
    for (int k = 0; k < 3; ++k) {
      if (!s.sameorder[f][k])
        for (int d = 0; d < nde; ++d) if (d % 2 == 1) ce[k][d] *= -1;
      int est = s.dof_start(1, s.num_edge_f[f][k]);
      for (int d = 0; d < nde; ++d) rhs[est + d] += ce[k][d];
    }
    int fst = s.dof_start(2, f);
    for (int d = 0; d < ndf; ++d) rhs[fst + d] += cf[d];
  }
}

// Dirichlet imposition with modal interpolation (src/Discretize.cpp:1778-1893)
// on the assembled CSR. uv[nv] vertex values, ue[ne][q1] values at edge
// quadrature points (start*(1-t)+end*t along num_node_e), uf[nf][q2] values at
// face quadrature points (bary of num_node_f). flag_boundary: mark to impose.
// Also writes the imposed modal values into bval (ndof, NaN elsewhere).
void or_dirichlet(void* p, const double* uv, const double* ue, const double* uf, int flag, int modify, double* rhs, double* bval) {
  OrHandle* h = (OrHandle*)p;
  Space& s = h->s;
  CSR& A = *h->A;
  const Quad& q = s.quad;
  int M = s.M, q1 = (int)q.t1.size(), q2 = (int)q.p2.size(), nde = s.ndg[1], ndf = s.ndg[2];
  if (bval) for (int i = 0; i < s.ndof; ++i) bval[i] = NAN;
  for (int dim = 0; dim <= 2; ++dim)
    for (int g = 0; g < s.ngeo[dim]; ++g) {
      if (s.mark[dim][g] != flag) continue;
      int row0 = s.dof_start(dim, g);
      std::vector<double> val(s.ndg[dim], 0.0);
      if (dim == 0) val[0] = uv[g];
      if (dim == 1) {
        int ls = s.dof_start(0, s.num_node_e[g][0]), le = s.dof_start(0, s.num_node_e[g][1]);
        double cs = rhs[ls] / A.diag(ls), ce = rhs[le] / A.diag(le);
        for (int l = 2; l <= M; ++l) {
          double cnt = 0;
          for (int pq = 0; pq < q1; ++pq)
            cnt += q.w1[pq] * (ue[(size_t)g * q1 + pq] - cs * (1 - q.t1[pq]) - ce * q.t1[pq]) * s.interp1[(size_t)pq * nde + l - 2];
          val[l - 2] = cnt * -l * (2 * l - 1) / (2 * (l - 1));
        }
      }
      if (dim == 2) {
        double cvert[3];
        for (int k = 0; k < 3; ++k) { int pos = s.dof_start(0, s.num_node_f[g][k]); cvert[k] = rhs[pos] / A.diag(pos); }
        std::vector<std::vector<double>> ved(3, std::vector<double>(nde));
        for (int k = 0; k < 3; ++k) {
          int est = s.dof_start(1, s.num_edge_f[g][k]);
          for (int l = 2; l <= M; ++l) {
            ved[k][l - 2] = rhs[est + l - 2] / A.diag(est + l - 2);
            if (!s.sameorder[g][k] && l % 2 == 1) ved[k][l - 2] *= -1;
          }
        }
        for (int l1 = 2; l1 <= M; ++l1)
          for (int l2 = 1; l2 <= M - l1; ++l2) {
            int id = (l1 + l2 - 2) * (l1 + l2 - 3) / 2 + l2 - 1;
            double cnt = 0;
            for (int pq = 0; pq < q2; ++pq) {
              double xp = q.p2[pq][0], yp = q.p2[pq][1], rp = q.p2[pq][2];
              double cp = uf[(size_t)g * q2 + pq];
              for (int k = 0; k < 3; ++k) cp -= cvert[k] * q.p2[pq][(k + 2) % 3];
              for (int l = 2; l <= M; ++l) {
                const double* ad = &s.addv[((size_t)pq * nde + l - 2) * 2];
                cp += 2 * (xp * yp * ved[0][l - 2] * ad[1] + yp * rp * ved[1][l - 2] * ad[1] + xp * rp * ved[2][l - 2] * ad[0]);
              }
              cnt += q.w2[pq] * cp * s.interp2[(size_t)pq * ndf + id];
            }
            val[id] = cnt * -l1 * (2 * l1 - 1) * (2 * l1 + 2 * l2 - 1) / (std::pow(2, l1) * (l1 - 1));
          }
      }
      for (int i = 0; i < s.ndg[dim]; ++i) {
        int row = row0 + i;
        if (bval) bval[row] = val[i];
        for (long k = A.rs[row]; k < A.rs[row + 1]; ++k) {
          int c = A.col[k];
          if (c == row) rhs[row] = A.diag(row) * val[i];
          else {
            long pc = A.find(c, row);
            rhs[c] -= A.v[pc] * val[i];
            if (modify) { A.v[k] = 0; A.v[pc] = 0; }
          }
        }
      }
    }
}

// impose_zero_boundary_condition on matrix and/or vector for mark==flag
void or_zero_bc(void* p, int flag, int on_matrix, double* vec) {
  OrHandle* h = (OrHandle*)p;
  Space& s = h->s;
  for (int dim = 0; dim <= 2; ++dim)
    for (int g = 0; g < s.ngeo[dim]; ++g) {
      if (s.mark[dim][g] != flag) continue;
      int row0 = s.dof_start(dim, g);
      for (int i = 0; i < s.ndg[dim]; ++i) {
        int row = row0 + i;
        if (vec) vec[row] = 0;
        if (on_matrix)
          for (long k = h->A->rs[row] + 1; k < h->A->rs[row + 1]; ++k) {
            int c = h->A->col[k];
            h->A->v[k] = 0;
            h->A->v[h->A->find(c, row)] = 0;
          }
      }
    }
}

// calc_sum_dof (src/Miscellaneous.cpp:557-587)
void or_sum_dof(void* p, int* sd) {
  Space& s = ((OrHandle*)p)->s;
  int M = s.M;
  for (int g = 0; g < s.nv; ++g) sd[s.dof_start(0, g)] = 1;
  for (int g = 0; g < s.ne; ++g) {
    int pos = s.dof_start(1, g);
    for (int l = 2; l <= M; ++l) sd[pos + l - 2] = l;
  }
  for (int g = 0; g < s.nf; ++g) {
    int pos = s.dof_start(2, g);
    for (int l1 = 2; l1 <= M; ++l1)
      for (int l2 = 1; l2 <= M - l1; ++l2) sd[pos + (1 + l1 + l2 - 3) * (l1 + l2 - 3) / 2 + l2 - 1] = l1 + l2;
  }
  for (int g = 0; g < s.nt; ++g) {
    int pos = s.dof_start(3, g);
    for (int l1 = 2; l1 <= M; ++l1)
      for (int l2 = 1; l2 <= M - l1; ++l2)
        for (int l3 = 1; l3 <= M - l1 - l2; ++l3) sd[pos + s.corr.i2n(l1 - 2, l2 - 1, l3 - 1)] = l1 + l2 + l3;
  }
}

}  // extern "C"

// ============================================================================
// Solvers (src/TSEMSolver.cpp)
// ============================================================================
namespace oracle {

struct PMG {
  int n_project = 0;
  std::vector<std::vector<int>> core_at;  // per level: fine position -> coarse index or -1
  std::vector<CSR> level;                 // level[0] aliases the finest (copy)
  std::vector<int> n_dof_layer;
};

static void gs_forward(const CSR& A, double* x, const double* r, int steps) {
  for (int it = 0; it < steps; ++it)
    for (int row = 0; row < A.n; ++row) {
      double r0 = r[row];
      for (long k = A.rs[row] + 1; k < A.rs[row + 1]; ++k) r0 -= x[A.col[k]] * A.v[k];
      x[row] = r0 / A.v[A.rs[row]];
    }
}
static void gs_backward(const CSR& A, double* x, const double* r, int steps) {
  for (int it = 0; it < steps; ++it)
    for (int row = A.n - 1; row >= 0; --row) {
      double r0 = r[row];
      for (long k = A.rs[row] + 1; k < A.rs[row + 1]; ++k) r0 -= x[A.col[k]] * A.v[k];
      x[row] = r0 / A.v[A.rs[row]];
    }
}
static void gs_symmetric(const CSR& A, double* x, const double* r, int steps) {
  for (int it = 0; it < steps; ++it) { gs_forward(A, x, r, 1); gs_backward(A, x, r, 1); }
}

static void init_pmg(PMG& P, const CSR& A, int p, const int* sum_dof) {
  int n = A.n;
  P.n_project = p - 1;
  std::vector<int> order(P.n_project + 1);
  order[0] = p;
  order[P.n_project] = 1;
  for (int i = 1; i < P.n_project; ++i) order[i] = order[i - 1] - 1;
  P.core_at.assign(P.n_project, {});
  P.n_dof_layer.assign(P.n_project + 1, 0);
  P.n_dof_layer[0] = n;
  std::vector<std::vector<int>> loc_inv(P.n_project, std::vector<int>(n, -1));
  int n_this = n;
  for (int ip = 0; ip < P.n_project; ++ip) {
    int pf = order[ip], pc = order[ip + 1];
    int nnext = 0, pa = -1;
    P.core_at[ip].assign(n_this, -1);
    for (int d = 0; d < n; ++d) {
      bool ff = sum_dof[d] <= pf, fc = sum_dof[d] <= pc;
      if (ff) pa++;
      if (fc) P.core_at[ip][pa] = nnext++;
      if (ff) loc_inv[ip][d] = pa;
    }
    n_this = nnext;
    P.n_dof_layer[ip + 1] = n_this;
  }
  P.level.clear();
  P.level.push_back(A);
  for (int ip = 0; ip < P.n_project; ++ip) {
    CSR L;
    L.n = P.n_dof_layer[ip + 1];
    std::vector<std::vector<std::pair<int, double>>> rows(L.n);
    for (int r = 0; r < n; ++r) {
      if (loc_inv[ip][r] < 0) continue;
      int cr = P.core_at[ip][loc_inv[ip][r]];
      if (cr < 0) continue;
      for (long k = A.rs[r]; k < A.rs[r + 1]; ++k) {
        int c = A.col[k];
        if (loc_inv[ip][c] < 0) continue;
        int cc = P.core_at[ip][loc_inv[ip][c]];
        if (cc < 0) continue;
        rows[cr].push_back({cc, A.v[k]});
      }
    }
    L.rs.assign(L.n + 1, 0);
    for (int r = 0; r < L.n; ++r) {
      auto& R = rows[r];
      std::sort(R.begin(), R.end(), [&](const std::pair<int, double>& a, const std::pair<int, double>& b) {
        if ((a.first == r) != (b.first == r)) return a.first == r;
        return a.first < b.first;
      });
      L.rs[r + 1] = L.rs[r] + (long)R.size();
      for (auto& pr : R) { L.col.push_back(pr.first); L.v.push_back(pr.second); }
    }
    P.level.push_back(std::move(L));
  }
}

// ite_PMG (src/TSEMSolver.cpp:1119-1184); the Gauss-Seidel sweeps are
// sequential as in the reference, the level residuals' products use nthreads
static void ite_pmg(const PMG& P, double* x, const double* r, int ite_step, int smooth, int nthreads = 1) {
  int np = P.n_project;
  std::vector<std::vector<double>> px(np + 1), pr(np + 1);
  px[0].assign(x, x + P.level[0].n);
  pr[0].assign(r, r + P.level[0].n);
  for (int i = 1; i <= np; ++i) { px[i].assign(P.level[i].n, 0.0); pr[i].assign(P.level[i].n, 0.0); }
  for (int it = 0; it < ite_step; ++it) {
    for (int i = 0; i < np; ++i) {
      gs_forward(P.level[i], px[i].data(), pr[i].data(), smooth);
      std::vector<double> v0(P.level[i].n), v1(pr[i]);
      P.level[i].vmult(px[i].data(), v0.data(), nthreads);
      for (int k = 0; k < P.level[i].n; ++k) v1[k] -= v0[k];
      for (int k = 0; k < P.level[i].n; ++k)
        if (P.core_at[i][k] >= 0) pr[i + 1][P.core_at[i][k]] = v1[k];
      std::fill(px[i + 1].begin(), px[i + 1].end(), 0.0);
    }
    gs_symmetric(P.level[np], px[np].data(), pr[np].data(), (smooth + 1) / 2);
    for (int i = np - 1; i >= 0; --i) {
      for (int k = 0; k < P.level[i].n; ++k)
        if (P.core_at[i][k] >= 0) px[i][k] += px[i + 1][P.core_at[i][k]];
      gs_backward(P.level[i], px[i].data(), pr[i].data(), smooth);
    }
  }
  std::copy(px[0].begin(), px[0].end(), x);
}

}  // namespace oracle

extern "C" {

// solve_PCG (src/TSEMSolver.cpp:1775-1930) on the handle's CSR.
// precond: 1 Jacobi, 4 p-MG (needs sum_dof, p, ite_step, smooth).
// criterion: 0 = reference (l1 norm of recursive residual < tol);
//            1 = relative ||r||_2/||b||_2 < tol on the recursive residual.
// nthreads > 1: the products and vector sweeps run on that many OpenMP
// threads (the sums then run in another order; the Gauss-Seidel sweeps stay
// sequential, as in the reference); 1 is the sequential reference form.
// stats out: [iterations, final l1 residual, final rel l2 residual,
//             p-MG setup seconds, solve seconds] (steady clock)
int or_pcg_mt(void* p, double* x, const double* b, double tol, int max_step, int precond, int ite_step, int smooth,
              const int* sum_dof, int criterion, int nthreads, double* stats) {
  OrHandle* h = (OrHandle*)p;
  const CSR& A = *h->A;
  const int n = A.n;
  const int nt = nthreads > 1 ? nthreads : 1;
  auto t0 = std::chrono::steady_clock::now();
  PMG P;
  if (precond == 4) init_pmg(P, A, h->s.M, sum_dof);
  auto t1 = std::chrono::steady_clock::now();
  std::vector<double> tmp(n), rk(n), yk(x, x + n), pk(b, b + n);
  A.vmult(x, rk.data(), nt);
  for (int i = 0; i < n; ++i) rk[i] -= b[i];
  auto precondition = [&](std::vector<double>& y) {
    if (precond == 1) {
#pragma omp parallel for schedule(static) num_threads(nt) if (nt > 1)
      for (int i = 0; i < n; ++i) y[i] = rk[i] / A.diag(i);
    } else if (precond == 4) {
      ite_pmg(P, y.data(), rk.data(), ite_step, smooth, nt);
    } else {
      y = rk;
    }
  };
  auto dot = [&](const std::vector<double>& a, const std::vector<double>& c) {
    double s = 0;
#pragma omp parallel for schedule(static) reduction(+ : s) num_threads(nt) if (nt > 1)
    for (int i = 0; i < n; ++i) s += a[i] * c[i];
    return s;
  };
  precondition(yk);
  for (int i = 0; i < n; ++i) pk[i] = -yk[i];
  double bnorm = 0;
  for (int i = 0; i < n; ++i) bnorm += b[i] * b[i];
  bnorm = std::sqrt(bnorm);
  auto resid = [&]() {
    double s = 0;
    if (criterion == 0) {
#pragma omp parallel for schedule(static) reduction(+ : s) num_threads(nt) if (nt > 1)
      for (int i = 0; i < n; ++i) s += std::fabs(rk[i]);
      return s;
    }
    return std::sqrt(dot(rk, rk)) / bnorm;
  };
  double residual = resid();
  int iter = 0;
  double rty = dot(rk, yk);
  while (residual >= tol) {
    A.vmult(pk.data(), tmp.data(), nt);
    double pap = dot(pk, tmp);
    double alpha = rty / pap;
#pragma omp parallel for schedule(static) num_threads(nt) if (nt > 1)
    for (int i = 0; i < n; ++i) {
      x[i] += alpha * pk[i];
      rk[i] += alpha * tmp[i];
      yk[i] = 0.0;
    }
    precondition(yk);
    double rty1 = dot(rk, yk);
    double beta = rty1 / rty;
#pragma omp parallel for schedule(static) num_threads(nt) if (nt > 1)
    for (int i = 0; i < n; ++i) pk[i] = beta * pk[i] - yk[i];
    residual = resid();
    iter++;
    if (iter > max_step) break;
    rty = rty1;
  }
  auto t2 = std::chrono::steady_clock::now();
  double l1 = 0, l2 = 0;
  for (int i = 0; i < n; ++i) { l1 += std::fabs(rk[i]); l2 += rk[i] * rk[i]; }
  stats[0] = iter; stats[1] = l1; stats[2] = std::sqrt(l2) / bnorm;
  stats[3] = std::chrono::duration<double>(t1 - t0).count();
  stats[4] = std::chrono::duration<double>(t2 - t1).count();
  return residual < tol ? 0 : 1;
}

int or_pcg(void* p, double* x, const double* b, double tol, int max_step, int precond, int ite_step, int smooth,
           const int* sum_dof, int criterion, double* stats) {
  double st[5];
  int rc = or_pcg_mt(p, x, b, tol, max_step, precond, ite_step, smooth, sum_dof, criterion, 1, st);
  for (int k = 0; k < 3; ++k) stats[k] = st[k];
  return rc;
}

}  // extern "C"
