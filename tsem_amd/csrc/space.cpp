// Host-side space setup (see space.hpp).
#include "space.hpp"

#include <algorithm>
#include <cmath>
#include <fstream>
#include <numeric>
#include <sstream>
#include <stdexcept>
#include <unordered_map>

namespace tsem {

static const int FP[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
static const int EP[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
// element edge opposite each vertex of element face k (face vertex order FP[k])
static const int FE[4][3] = {{5, 4, 3}, {5, 2, 1}, {4, 2, 0}, {3, 1, 0}};

void FaceTable::load(const std::string& path) {
  std::ifstream in(path);
  if (!in) throw std::runtime_error("cannot open face table " + path);
  std::string t1, t2;
  in >> t1 >> mmax >> t2 >> nface;
  terms.assign(6 * nface, {});
  for (int r = 0; r < 6 * nface; ++r) {
    int v, d;
    in >> v >> d;
    if (v < 0 || v > 5 || d < 0 || d >= nface) throw std::runtime_error("bad face table row");
    auto& T = terms[v * nface + d];
    for (int ie = 0; ie < 4; ++ie) {
      int cnt;
      in >> cnt;
      for (int c = 0; c < cnt; ++c) {
        Term tm;
        tm.ie = ie;
        in >> tm.idx >> tm.w;
        T.push_back(tm);
      }
    }
  }
  if (!in) throw std::runtime_error("truncated face table");
}

static inline uint64_t key2(int a, int b) {
  if (a > b) std::swap(a, b);
  return ((uint64_t)a << 32) | (uint32_t)b;
}
static inline uint64_t key3(int a, int b, int c) {
  int s[3] = {a, b, c};
  std::sort(s, s + 3);
  return ((uint64_t)s[0] << 42) | ((uint64_t)s[1] << 21) | (uint64_t)s[2];
}

void build_topology(const int32_t* tet, int nt, int nv, Topology& T) {
  if (nv >= (1 << 21)) throw std::runtime_error("mesh too large for face keys");
  T.nv = nv; T.nt = nt;
  T.edge_vtx.clear(); T.face_vtx.clear(); T.face_edge.clear();
  T.tet_face.assign((size_t)nt * 4, -1);
  T.tet_edge.assign((size_t)nt * 6, -1);
  std::unordered_map<uint64_t, int32_t> em, fm;
  em.reserve((size_t)nt * 2);
  fm.reserve((size_t)nt * 3);
  for (int t = 0; t < nt; ++t) {
    const int32_t* v = tet + 4 * (size_t)t;
    for (int k = 0; k < 6; ++k) {
      int a = v[EP[k][0]], b = v[EP[k][1]];
      auto ins = em.emplace(key2(a, b), (int32_t)(T.edge_vtx.size() / 2));
      if (ins.second) { T.edge_vtx.push_back(a); T.edge_vtx.push_back(b); }
      T.tet_edge[6 * (size_t)t + k] = ins.first->second;
    }
  }
  T.ne = (int)(T.edge_vtx.size() / 2);
  for (int t = 0; t < nt; ++t) {
    const int32_t* v = tet + 4 * (size_t)t;
    for (int k = 0; k < 4; ++k) {
      int a = v[FP[k][0]], b = v[FP[k][1]], c = v[FP[k][2]];
      auto ins = fm.emplace(key3(a, b, c), (int32_t)(T.face_vtx.size() / 3));
      if (ins.second) {
        T.face_vtx.push_back(a); T.face_vtx.push_back(b); T.face_vtx.push_back(c);
        T.face_edge.push_back(T.tet_edge[6 * (size_t)t + FE[k][0]]);
        T.face_edge.push_back(T.tet_edge[6 * (size_t)t + FE[k][1]]);
        T.face_edge.push_back(T.tet_edge[6 * (size_t)t + FE[k][2]]);
      }
      T.tet_face[4 * (size_t)t + k] = ins.first->second;
    }
  }
  T.nf = (int)(T.face_vtx.size() / 3);
}

void canonical_femdof(const Topology& T, const int32_t* tet, std::vector<int32_t>& w) {
  w.assign((size_t)T.nv + T.ne, -1);
  int32_t c = 0;
  for (int t = 0; t < T.nt; ++t) {
    for (int k = 0; k < 4; ++k) {
      int32_t& x = w[tet[4 * (size_t)t + k]];
      if (x < 0) x = c++;
    }
    for (int k = 0; k < 6; ++k) {
      int32_t& x = w[(size_t)T.nv + T.tet_edge[6 * (size_t)t + k]];
      if (x < 0) x = c++;
    }
  }
}

int64_t Space::geo_dof(int dim, int i) const {
  size_t off = 0;
  if (dim > 0) off += topo.nv;
  if (dim > 1) off += topo.ne;
  if (dim > 2) off += topo.nf;
  return dof_start[off + i];
}

// Volumes and geometric factors of every element from vertex coordinates
// (the mesh connectivity stays): pair weights 2/(3V) x dihedral factor
// (src/Discretize.cpp:1352-1370), the metric Ghat and 6V.
void Space::set_geometry(const double* xyz_new) {
  const int nt = topo.nt;
  if (xyz_new != xyz.data()) xyz.assign(xyz_new, xyz_new + xyz.size());
  vol.resize(nt);
  for (int t = 0; t < nt; ++t) {
    const double* p0 = &xyz[3 * (size_t)tet[4 * t]];
    double a[3][3];
    for (int r = 1; r < 4; ++r)
      for (int d = 0; d < 3; ++d) a[r - 1][d] = xyz[3 * (size_t)tet[4 * t + r] + d] - p0[d];
    double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) - a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                 a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
    vol[t] = std::fabs(det) / 6.0;
  }
  pairw.assign((size_t)nt * 6, 0.0);
  ghat.assign((size_t)nt * 6, 0.0);
  sixv.assign(nt, 0.0);
  for (int t = 0; t < nt; ++t) {
    double* w = &pairw[6 * (size_t)t];
    double coef = 2.0 / (3 * vol[t]);
    for (int i1 = 0; i1 <= 2; ++i1)
      for (int i2 = i1 + 1; i2 <= 3; ++i2) {
        int ip1 = (i1 == 0) ? (i2 % 2 + 1) : 0;
        int ip2 = 6 - i1 - i2 - ip1;
        const double* p1 = &xyz[3 * (size_t)tet[4 * t + i1]];
        const double* p2 = &xyz[3 * (size_t)tet[4 * t + i2]];
        const double* ps = &xyz[3 * (size_t)tet[4 * t + ip1]];
        const double* pe = &xyz[3 * (size_t)tet[4 * t + ip2]];
        double a[3], b[3], c[3];
        for (int d = 0; d < 3; ++d) { a[d] = pe[d] - ps[d]; b[d] = p1[d] - ps[d]; c[d] = p2[d] - ps[d]; }
        double u[3] = {a[1] * b[2] - a[2] * b[1], -a[0] * b[2] + a[2] * b[0], a[0] * b[1] - a[1] * b[0]};
        double vv[3] = {a[1] * c[2] - a[2] * c[1], -a[0] * c[2] + a[2] * c[0], a[0] * c[1] - a[1] * c[0]};
        double dih = 0.25 * (u[0] * vv[0] + u[1] * vv[1] + u[2] * vv[2]);
        if (std::fabs(dih) < 1e-12) continue;
        // pair (0,j): d_j ; pair (i,j): d_j - d_i ; slots as the reference's
        // multi-index numbering minus 4: (200)0 (110)1 (101)2 (020)3 (011)4 (002)5
        int slot;
        if (i1 == 0) slot = (i2 == 1) ? 0 : (i2 == 2) ? 3 : 5;
        else slot = (i1 == 1 && i2 == 2) ? 1 : (i1 == 1 && i2 == 3) ? 2 : 4;
        w[slot] += coef * dih;
      }
    double* g = &ghat[6 * (size_t)t];
    g[0] = w[0] + w[1] + w[2];
    g[1] = w[1] + w[3] + w[4];
    g[2] = w[2] + w[4] + w[5];
    g[3] = -w[1];
    g[4] = -w[2];
    g[5] = -w[4];
    sixv[t] = 6 * vol[t];
  }
}

void Space::build(int M_, int nv, const double* xyz_, int nt, const int32_t* tet_, const int32_t* femdof_w,
                  const std::string& table_path) {
  M = M_;
  ref = RefElement::cached(M);
  K = ref.K;
  table.load(table_path);
  if (table.mmax < M) throw std::runtime_error("face table does not cover this order");
  xyz.assign(xyz_, xyz_ + 3 * (size_t)nv);
  tet.assign(tet_, tet_ + 4 * (size_t)nt);
  for (size_t i = 0; i < tet.size(); ++i)
    if (tet[i] < 0 || tet[i] >= nv) throw std::runtime_error("tet vertex index out of range");
  build_topology(tet.data(), nt, nv, topo);
  if (femdof_w) femdof.assign(femdof_w, femdof_w + (size_t)nv + topo.ne);
  else canonical_femdof(topo, tet.data(), femdof);
  for (int d = 0; d < 4; ++d) {
    // C(M-1, d), zero when M-1 < d
    int n = M - 1, v = 1;
    if (n < d) v = 0;
    else for (int i = 0; i < d; ++i) v = v * (n - i) / (i + 1);
    ndg[d] = v;
  }
  const int ne = topo.ne, nf = topo.nf;

  // ---- first-visit orientation ----
  edge_node.assign(ne, {-1, -1});
  face_node.assign(nf, {-1, -1, -1});
  face_edges.assign(nf, {-1, -1, -1});
  face_rev.assign(nf, 0);
  tet_eflip.assign(nt, 0);
  tet_fv.assign(nt, {0, 0, 0, 0});
  std::vector<uint8_t> seen_e(ne, 0), seen_f(nf, 0);
  for (int t = 0; t < nt; ++t) {
    const int32_t* v = &tet[4 * (size_t)t];
    const int32_t* te = &topo.tet_edge[6 * (size_t)t];
    for (int k = 0; k < 6; ++k) {
      int ge = te[k];
      if (!seen_e[ge]) { edge_node[ge] = {v[EP[k][0]], v[EP[k][1]]}; seen_e[ge] = 1; }
      else if (edge_node[ge][0] != v[EP[k][0]]) tet_eflip[t] |= (uint8_t)(1u << k);
    }
    for (int k = 0; k < 4; ++k) {
      int gf = topo.tet_face[4 * (size_t)t + k];
      if (!seen_f[gf]) {
        seen_f[gf] = 1;
        for (int j = 0; j < 3; ++j) { face_node[gf][j] = v[FP[k][j]]; face_edges[gf][j] = te[FE[k][j]]; }
        // edge j of the face is opposite face vertex j; it runs "in order"
        // when it starts at the lower-ordered remaining vertex
        if (edge_node[face_edges[gf][0]][0] != v[FP[k][1]]) face_rev[gf] |= 1;
        if (edge_node[face_edges[gf][1]][0] != v[FP[k][0]]) face_rev[gf] |= 2;
        if (edge_node[face_edges[gf][2]][0] != v[FP[k][0]]) face_rev[gf] |= 4;
        tet_fv[t][k] = 0;
      } else {
        int ord[3] = {-1, -1, -1};
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b)
            if (v[FP[k][a]] == face_node[gf][b]) ord[a] = b;
        int code = ord[0] * 100 + ord[1] * 10 + ord[2];
        int var;
        switch (code) {
          case 12: var = 0; break;
          case 21: var = 1; break;
          case 102: var = 2; break;
          case 120: var = 4; break;  // inverse of the cyclic shift 1,2,0
          case 201: var = 3; break;
          case 210: var = 5; break;
          default: throw std::runtime_error("inconsistent face orientation");
        }
        tet_fv[t][k] = (uint8_t)var;
      }
    }
  }
  // ---- global DOF layout: stable order of all geometries by exact keys ----
  size_t ng = (size_t)nv + ne + nf + nt;
  std::vector<int64_t> key(ng);
  for (int i = 0; i < nv; ++i) key[i] = 12 * (int64_t)femdof[i];
  for (int i = 0; i < ne; ++i) key[nv + i] = 12 * (int64_t)femdof[(size_t)nv + i];
  for (int i = 0; i < nf; ++i) {
    int64_t s = 0;
    for (int j = 0; j < 3; ++j) s += femdof[topo.face_vtx[3 * (size_t)i + j]];
    key[(size_t)nv + ne + i] = 4 * s;
  }
  for (int i = 0; i < nt; ++i) {
    int64_t s = 0;
    for (int j = 0; j < 4; ++j) s += femdof[tet[4 * (size_t)i + j]];
    key[(size_t)nv + ne + nf + i] = 3 * s;
  }
  std::vector<int32_t> order(ng);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return key[a] < key[b]; });
  auto dim_of = [&](size_t g) { return g < (size_t)nv ? 0 : g < (size_t)nv + ne ? 1 : g < (size_t)nv + ne + nf ? 2 : 3; };
  dof_start.assign(ng, 0);
  int64_t run = 0;
  for (size_t r = 0; r < ng; ++r) {
    dof_start[order[r]] = run;
    run += ndg[dim_of(order[r])];
  }
  ndof = run;
  if (ndof >= (1LL << 30)) throw std::runtime_error("too many DOFs for 30-bit indices");

  // ---- slot tables ----
  const ModeIndex& mi = ref.idx;
  slot_kind.assign(K, -1); slot_geo.assign(K, -1); slot_sub.assign(K, -1);
  for (int k = 0; k < 4; ++k) { slot_kind[k] = 0; slot_geo[k] = k; slot_sub[k] = 0; }
  for (int l = 2; l <= M; ++l) {
    int ids[6] = {mi.number(l, 0, 0), mi.number(0, l, 0), mi.number(0, 0, l),
                  mi.number(1, l - 1, 0), mi.number(1, 0, l - 1), mi.number(0, 1, l - 1)};
    for (int k = 0; k < 6; ++k) { slot_kind[ids[k]] = 1; slot_geo[ids[k]] = k; slot_sub[ids[k]] = l - 2; }
  }
  for (int l1 = 2; l1 <= M; ++l1)
    for (int l2 = 1; l1 + l2 <= M; ++l2) {
      int d = (l1 + l2 - 3) * (l1 + l2 - 2) / 2 + l2 - 1;
      int ids[4] = {mi.number(1, l1 - 1, l2), mi.number(0, l1, l2), mi.number(l1, 0, l2), mi.number(l1, l2, 0)};
      for (int k = 0; k < 4; ++k) { slot_kind[ids[k]] = 2; slot_geo[ids[k]] = k; slot_sub[ids[k]] = d; }
    }
  for (int d = 0; d < ndg[3]; ++d) {
    auto a = mi.mi[d];
    int id = mi.number(a[0] + 2, a[1] + 1, a[2] + 1);
    slot_kind[id] = 3; slot_geo[id] = 0; slot_sub[id] = d;
  }
  for (int k = 0; k < K; ++k)
    if (slot_kind[k] < 0) throw std::runtime_error("unassigned local mode");

  // ---- element descriptors ----
  desc.assign(nt, ElemDesc{});
  for (int t = 0; t < nt; ++t) {
    ElemDesc& D = desc[t];
    const int32_t* te = &topo.tet_edge[6 * (size_t)t];
    for (int k = 0; k < 4; ++k) D.v[k] = (int32_t)geo_dof(0, tet[4 * (size_t)t + k]);
    for (int k = 0; k < 6; ++k) {
      D.e[k] = (int32_t)geo_dof(1, te[k]);
      if (tet_eflip[t] & (1u << k)) D.e[k] |= (int32_t)0x80000000u;
    }
    for (int k = 0; k < 4; ++k) {
      int gf = topo.tet_face[4 * (size_t)t + k];
      D.f[k] = (int32_t)geo_dof(2, gf);
      int code = tet_fv[t][k];
      for (int ie = 0; ie < 3; ++ie) {
        int le = -1;
        for (int q = 0; q < 6; ++q) if (te[q] == face_edges[gf][ie]) le = q;
        if (le < 0) throw std::runtime_error("face edge not on element");
        code |= le << (3 + 3 * ie);
        if (face_rev[gf] & (1u << ie)) code |= 1 << (12 + ie);
      }
      D.fo[k] = code;
    }
    D.t = (int32_t)geo_dof(3, t);
  }

  // ---- DOF degrees (calc_sum_dof, src/Miscellaneous.cpp:557-587) ----
  sum_dof.assign(ndof, 0);
  for (int i = 0; i < nv; ++i) sum_dof[geo_dof(0, i)] = 1;
  for (int i = 0; i < ne; ++i)
    for (int l = 2; l <= M; ++l) sum_dof[geo_dof(1, i) + l - 2] = l;
  for (int i = 0; i < nf; ++i)
    for (int l1 = 2; l1 <= M; ++l1)
      for (int l2 = 1; l1 + l2 <= M; ++l2) sum_dof[geo_dof(2, i) + (l1 + l2 - 3) * (l1 + l2 - 2) / 2 + l2 - 1] = l1 + l2;
  for (int i = 0; i < nt; ++i)
    for (int d = 0; d < ndg[3]; ++d) {
      auto a = mi.mi[d];
      sum_dof[geo_dof(3, i) + d] = a[0] + a[1] + a[2] + 4;
    }

  set_geometry(xyz.data());

  // ---- transposed gather map for the deterministic scatter ----
  std::vector<std::pair<int64_t, double>> terms;
  std::vector<int64_t> cnt(ndof + 1, 0);
  for (int t = 0; t < nt; ++t)
    for (int i = 0; i < K; ++i) {
      terms.clear();
      gather_terms(t, i, terms);
      for (auto& pr : terms) cnt[pr.first + 1]++;
    }
  for (int64_t g = 0; g < ndof; ++g) cnt[g + 1] += cnt[g];
  sc_ptr = cnt;
  sc_code.assign(cnt[ndof], 0);
  sc_w.assign(cnt[ndof], 0.0);
  std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
  for (int t = 0; t < nt; ++t)
    for (int i = 0; i < K; ++i) {
      terms.clear();
      gather_terms(t, i, terms);
      int64_t slot = (int64_t)t * K + i;
      if (slot >= (1LL << 30)) throw std::runtime_error("too many element slots");
      for (auto& pr : terms) {
        int64_t pos = fill[pr.first]++;
        int32_t code = (int32_t)slot;
        if (pr.second == -1.0) code |= 1 << 30;
        else if (pr.second != 1.0) code |= (int32_t)0x80000000u;
        sc_code[pos] = code;
        sc_w[pos] = pr.second;
      }
    }
  for (int d = 0; d < 3; ++d) mark[d].clear();
}

void Space::gather_terms(int t, int i, std::vector<std::pair<int64_t, double>>& out) const {
  const ElemDesc& D = desc[t];
  int kind = slot_kind[i], geo = slot_geo[i], sub = slot_sub[i];
  if (kind == 0) { out.push_back({D.v[geo], 1.0}); return; }
  if (kind == 1) {
    int l = sub + 2;
    bool flip = (D.e[geo] & 0x80000000) != 0;
    out.push_back({(int64_t)(D.e[geo] & 0x7fffffff) + sub, (flip && (l % 2 == 1)) ? -1.0 : 1.0});
    return;
  }
  if (kind == 3) { out.push_back({(int64_t)D.t + sub, 1.0}); return; }
  int code = D.fo[geo], var = code & 7;
  for (auto& tm : table.terms[var * table.nface + sub]) {
    if (tm.ie < 3) {
      int le = (code >> (3 + 3 * tm.ie)) & 7;
      bool rev = (code >> (12 + tm.ie)) & 1;
      double w = tm.w;
      if (rev && tm.idx % 2 == 1) w = -w;
      out.push_back({(int64_t)(D.e[le] & 0x7fffffff) + tm.idx, w});
    } else {
      out.push_back({(int64_t)D.f[geo] + tm.idx, tm.w});
    }
  }
}

void Space::element_matrix(int t, double stiff, double mass_coef, std::vector<double>& Ke) const {
  Ke.assign((size_t)K * K, 0.0);
  const double* g = &ghat[6 * (size_t)t];
  const double G[9] = {g[0], g[3], g[4], g[3], g[1], g[5], g[4], g[5], g[2]};
  for (int ab = 0; ab < 9; ++ab) {
    double c = stiff * G[ab];
    if (c == 0) continue;
    for (size_t s = 0; s < (size_t)K * K; ++s) Ke[s] += c * ref.grad2[ab][s];
  }
  if (mass_coef != 0)
    for (size_t s = 0; s < (size_t)K * K; ++s) Ke[s] += mass_coef * sixv[t] * ref.mass[s];
}

void Space::diagonal(double stiff, double mass_coef, std::vector<double>& diag) const {
  diag.assign(ndof, 0.0);
  std::vector<std::pair<int64_t, double>> terms;
  struct E { int64_t g; int i; double w; };
  std::vector<E> all;
  for (int t = 0; t < topo.nt; ++t) {
    all.clear();
    for (int i = 0; i < K; ++i) {
      terms.clear();
      gather_terms(t, i, terms);
      for (auto& pr : terms) all.push_back({pr.first, i, pr.second});
    }
    std::sort(all.begin(), all.end(), [](const E& a, const E& b) { return a.g < b.g || (a.g == b.g && a.i < b.i); });
    const double* g = &ghat[6 * (size_t)t];
    const double G[9] = {g[0], g[3], g[4], g[3], g[1], g[5], g[4], g[5], g[2]};
    auto kij = [&](int i, int j) {
      double s = 0;
      size_t o = (size_t)i * K + j;
      for (int ab = 0; ab < 9; ++ab) s += stiff * G[ab] * ref.grad2[ab][o];
      if (mass_coef != 0) s += mass_coef * sixv[t] * ref.mass[o];
      return s;
    };
    for (size_t a = 0; a < all.size();) {
      size_t b = a;
      while (b < all.size() && all[b].g == all[a].g) ++b;
      double s = 0;
      for (size_t u = a; u < b; ++u)
        for (size_t v = a; v < b; ++v) s += all[u].w * all[v].w * kij(all[u].i, all[v].i);
      diag[all[a].g] += s;
      a = b;
    }
  }
}

}  // namespace tsem
