// Entity-slab tables (layout described in slab.hpp).  The transposed face
// rows restate the face re-orientation table of the gather
// (src/BuildSpace.cpp:1703-1771 applies it forward) with its terms grouped
// by the global DOF they reach.
#include "slab.hpp"

#include <algorithm>
#include <stdexcept>

namespace tsem {

namespace {
// Slab entries of one element on the host.  Y holds the element's results
// in slab order: row 4 + e n1 + l is edge e's DOF l, row 4 + 6 n1 + f n2 + d
// face f's mode d, vertices first, interior last (Slab::perm).
struct HostTab {
  int n1, n2, JT;
  const uint16_t* tk;
  const double* tw;
};
// sum_t w_t Y[fb + d_t] over transposed row `row` (fb: first Y row of the face)
double slab_row(const HostTab& B, const double* Y, int fb, int row) {
  const uint16_t* k = B.tk + (size_t)row * B.JT;
  const double* w = B.tw + (size_t)row * B.JT;
  const int nterm = k[0] >> 8;
  double v = 0.0;
  for (int t = 0; t < nterm; ++t) v += w[t] * Y[fb + (k[t] & 255)];
  return v;
}
// edge e (E01, E02, E03, E12, E13, E23), DOF l: the edge mode with the edge
// sign, then the face modes of the two element faces holding the edge (face
// k is opposite vertex k), lower face first
double slab_edge(const HostTab& B, const double* Y, int e, int l, int32_t ow, int32_t ow2) {
  double v = Y[4 + e * B.n1 + l];
  if (((ow >> (24 + e)) & 1) && (l & 1)) v = -v;
  static const int FH[6][2] = {{2, 3}, {1, 3}, {1, 2}, {0, 3}, {0, 2}, {0, 1}};
  double part[2];
  for (int h = 0; h < 2; ++h) {
    const int f = FH[e][h], var = (ow >> (6 * f)) & 7, q = (ow2 >> (4 * e + 2 * h)) & 3;
    const double p = slab_row(B, Y, 4 + 6 * B.n1 + f * B.n2, 6 * B.n2 + (var * 3 + q) * B.n1 + l);
    part[h] = (((ow >> (6 * f + 3 + q)) & 1) && (l & 1)) ? -p : p;
  }
  return (v + part[0]) + part[1];
}
// face f, DOF i: the face modes that the face re-expression sends to it
double slab_face(const HostTab& B, const double* Y, int f, int i, int32_t ow) {
  return slab_row(B, Y, 4 + 6 * B.n1 + f * B.n2, ((ow >> (6 * f)) & 7) * B.n2 + i);
}

// (dimension, first global DOF) of entity j of an element: v0..3, e0..5, f0..3, interior
inline void entity(const ElemDesc& D, int j, int& dim, int64_t& gs) {
  if (j < 4) { dim = 0; gs = D.v[j]; }
  else if (j < 10) { dim = 1; gs = D.e[j - 4] & 0x7fffffff; }
  else if (j < 14) { dim = 2; gs = D.f[j - 10]; }
  else { dim = 3; gs = D.t; }
}
}  // namespace

void build_slab(const Space& S, Slab& out) {
  out = Slab{};
  const int nt = S.topo.nt;
  const int64_t ndof = S.ndof;
  const int n[4] = {S.ndg[0], S.ndg[1], S.ndg[2], S.ndg[3]};
  out.n1 = n[1]; out.n2 = n[2]; out.n3 = n[3];
  if ((int64_t)nt * S.K >= (1LL << 30)) { out.why = "slab larger than 2^30 entries"; return; }

  // incident elements per geometry, kept at the geometry's first DOF, then
  // spread over its DOFs
  std::vector<int32_t> m(ndof, 0), dimof(ndof, 0);
  for (int t = 0; t < nt; ++t)
    for (int j = 0; j < 15; ++j) {
      int dim; int64_t gs;
      entity(S.desc[t], j, dim, gs);
      if (n[dim] > 0) m[gs]++;
    }
  const int ng[4] = {S.topo.nv, S.topo.ne, S.topo.nf, nt};
  size_t gbase = 0;
  for (int dim = 0; dim < 4; ++dim) {
    for (int i = 0; i < ng[dim]; ++i) {
      int64_t gs = S.dof_start[gbase + i];
      for (int k = 0; k < n[dim]; ++k) { m[gs + k] = m[gs]; dimof[gs + k] = dim; }
    }
    gbase += ng[dim];
  }
  int mmax = 0;
  std::vector<int64_t> P(ndof + 1, 0);
  for (int64_t g = 0; g < ndof; ++g) { P[g + 1] = P[g] + m[g]; mmax = std::max(mmax, m[g]); }
  if (mmax >= 1024) { out.why = "a geometry with 1024 or more elements"; return; }
  out.total = P[ndof];
  if (out.total != (int64_t)nt * S.K) { out.why = "slab size differs from nt K"; return; }

  // per element: entity offsets by rank in element order, orientation words
  out.eoff.assign((size_t)nt * 17, 0);
  std::vector<int32_t> rank(ndof, 0);
  for (int t = 0; t < nt; ++t) {
    const ElemDesc& D = S.desc[t];
    int32_t* eo = &out.eoff[(size_t)t * 17];
    for (int j = 0; j < 15; ++j) {
      int dim; int64_t gs;
      entity(D, j, dim, gs);
      if (n[dim] == 0) continue;
      eo[j] = (int32_t)(P[gs] + (int64_t)rank[gs]++ * n[dim]);
    }
    int32_t w = 0, w2 = 0;
    for (int f = 0; f < 4; ++f) {
      w |= (D.fo[f] & 7) << (6 * f);
      w |= ((D.fo[f] >> 12) & 7) << (6 * f + 3);
    }
    for (int k = 0; k < 6; ++k)
      if (D.e[k] < 0) w |= 1 << (24 + k);
    // for each edge and each of its two faces (lower first; face k is
    // opposite vertex k): which edge of the face it is
    static const int EP[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
    for (int e = 0; e < 6; ++e) {
      int h = 0;
      for (int f = 0; f < 4; ++f) {
        if (f == EP[e][0] || f == EP[e][1]) continue;
        int q = -1;
        for (int ie = 0; ie < 3; ++ie)
          if (((D.fo[f] >> (3 + 3 * ie)) & 7) == e) q = ie;
        if (q < 0) { out.why = "an element face without one of its edges"; return; }
        w2 |= q << (4 * e + 2 * h);
        ++h;
      }
    }
    eo[15] = w;
    eo[16] = w2;
  }

  // per DOF: offset in its chunk's span, count, dimension
  const int64_t nch = (ndof + SLAB_ROWS - 1) / SLAB_ROWS;
  out.chunk.assign(2 * ((nch + SLAB_CHUNK_GROUP - 1) / SLAB_CHUNK_GROUP * SLAB_CHUNK_GROUP), 0);
  out.meta.assign(ndof, 0);
  std::vector<int64_t> first(ndof);
  for (int64_t g = 0; g < ndof; ++g) first[g] = -1;
  gbase = 0;
  for (int dim = 0; dim < 4; ++dim) {
    for (int i = 0; i < ng[dim]; ++i) {
      int64_t gs = S.dof_start[gbase + i];
      for (int k = 0; k < n[dim]; ++k) first[gs + k] = P[gs] + k;
    }
    gbase += ng[dim];
  }
  for (int64_t c = 0; c < nch; ++c) {
    const int64_t g0 = c * SLAB_ROWS, g1 = std::min(ndof, g0 + SLAB_ROWS);
    int64_t lo = -1, hi = -1;
    for (int64_t g = g0; g < g1; ++g) {
      if (first[g] < 0) { out.why = "DOF outside every geometry"; return; }
      if (m[g] == 0) continue;
      int64_t end = first[g] + (int64_t)(m[g] - 1) * (dimof[g] == 0 ? 1 : n[dimof[g]]) + 1;
      lo = lo < 0 ? first[g] : std::min(lo, first[g]);
      hi = std::max(hi, end);
    }
    if (lo < 0) lo = hi = 0;
    if (hi - lo >= (1 << 20)) { out.why = "chunk span of 2^20 or more entries"; return; }
    out.chunk[2 * c] = (int32_t)lo;
    out.chunk[2 * c + 1] = (int32_t)(hi - lo);
    out.chunk_max = std::max(out.chunk_max, (int)(hi - lo));
    for (int64_t g = g0; g < g1; ++g) {
      int64_t rel = m[g] ? first[g] - lo : 0;
      out.meta[g] = (int32_t)((uint32_t)rel | ((uint32_t)m[g] << 20) | ((uint32_t)dimof[g] << 30));
    }
  }

  // slab row of each local mode
  out.perm.assign(S.K, 0);
  for (int i = 0; i < S.K; ++i) {
    int kind = S.slot_kind[i], geo = S.slot_geo[i], sub = S.slot_sub[i];
    out.perm[i] = kind == 0 ? geo : kind == 1 ? 4 + geo * n[1] + sub
                : kind == 2 ? 4 + 6 * n[1] + geo * n[2] + sub : 4 + 6 * n[1] + 4 * n[2] + sub;
  }

  // transposed face rows
  const int nrow = 6 * n[2] + 18 * n[1];
  std::vector<std::vector<std::pair<int, double>>> rows(nrow);
  for (int v = 0; v < 6; ++v)
    for (int d = 0; d < n[2]; ++d)
      for (const auto& tm : S.table.terms[(size_t)v * S.table.nface + d]) {
        if (tm.ie == 3) {
          if (tm.idx >= n[2]) { out.why = "face term beyond the face DOFs"; return; }
          rows[(size_t)v * n[2] + tm.idx].push_back({d, tm.w});
        } else {
          if (tm.idx >= n[1]) { out.why = "edge term beyond the edge DOFs"; return; }
          rows[6 * (size_t)n[2] + ((size_t)v * 3 + tm.ie) * n[1] + tm.idx].push_back({d, tm.w});
        }
      }
  // variant 0 (the face seen as its first visitor saw it) must be the
  // identity with no edge terms: the element kernels copy such faces
  for (int i = 0; i < n[2]; ++i) {
    const auto& r = rows[i];
    if (r.size() != 1 || r[0].first != i || r[0].second != 1.0) { out.why = "face variant 0 is not the identity"; return; }
  }
  for (int q = 0; q < 3; ++q)
    for (int l = 0; l < n[1]; ++l)
      if (!rows[6 * (size_t)n[2] + (size_t)q * n[1] + l].empty()) { out.why = "face variant 0 has edge terms"; return; }
  // (with more than 32 face DOFs only sign-free ones: the masks are 32 bits)
  for (int v = 0; v < 6; ++v) {
    bool dg = true;
    uint32_t sg = 0;
    for (int i = 0; i < n[2] && dg; ++i) {
      const auto& r = rows[(size_t)v * n[2] + i];
      dg = r.size() == 1 && r[0].first == i && (r[0].second == 1.0 || (r[0].second == -1.0 && i < 32));
      if (dg && r[0].second < 0) sg |= 1u << i;
    }
    for (int q = 0; q < 3 && dg; ++q)
      for (int l = 0; l < n[1]; ++l) dg = dg && rows[6 * (size_t)n[2] + ((size_t)v * 3 + q) * n[1] + l].empty();
    if (dg) { out.diag |= 1 << v; out.sgn[v] = sg; }
  }
  int JT = 1;
  for (auto& r : rows) JT = std::max(JT, (int)r.size());
  if (JT > SLAB_JMAX || n[2] > 255) { out.why = "transposed face rows too long"; return; }
  out.JT = JT;
  const int nr = slab_rows(n[1], n[2]);
  out.nrow = nr;
  out.tk.assign((size_t)nr * JT, 0);
  out.tw.assign((size_t)nr * JT, 0.0);
  for (int r = 0; r < nrow; ++r) {
    for (size_t j = 0; j < rows[r].size(); ++j) {
      out.tk[(size_t)r * JT + j] = (uint16_t)rows[r][j].first;
      out.tw[(size_t)r * JT + j] = rows[r][j].second;
    }
    out.tk[(size_t)r * JT] |= (uint16_t)(rows[r].size() << 8);
  }
  out.ok = true;
}

void slab_apply_host(const Space& S, const Slab& L, const double* yl, double* y) {
  if (!L.ok) throw std::logic_error("no slab layout: " + L.why);
  const HostTab H{L.n1, L.n2, L.JT, L.tk.data(), L.tw.data()};
  const int K = S.K, nt = S.topo.nt;
  std::vector<double> slab(L.total, 0.0), Y(K);
  for (int t = 0; t < nt; ++t) {
    const int32_t* eo = &L.eoff[(size_t)t * 17];
    for (int i = 0; i < K; ++i) Y[L.perm[i]] = yl[(size_t)t * K + i];
    for (int a = 0; a < 4; ++a) slab[eo[a]] = Y[a];
    for (int e = 0; e < 6; ++e)
      for (int l = 0; l < L.n1; ++l) slab[(size_t)eo[4 + e] + l] = slab_edge(H, Y.data(), e, l, eo[15], eo[16]);
    for (int f = 0; f < 4; ++f)
      for (int i = 0; i < L.n2; ++i) slab[(size_t)eo[10 + f] + i] = slab_face(H, Y.data(), f, i, eo[15]);
    for (int i = 0; i < L.n3; ++i) slab[(size_t)eo[14] + i] = Y[4 + 6 * L.n1 + 4 * L.n2 + i];
  }
  for (int64_t g = 0; g < S.ndof; ++g) {
    const uint32_t meta = (uint32_t)L.meta[g];
    const int rel = meta & 0xfffff, m = (meta >> 20) & 0x3ff, dim = meta >> 30;
    const int stride = dim == 1 ? L.n1 : dim == 2 ? L.n2 : 1;
    const int64_t base = L.chunk[2 * (g / SLAB_ROWS)] + (int64_t)rel;
    double acc = 0.0;
    for (int k = 0; k < m; ++k) acc += slab[base + (int64_t)k * stride];
    y[g] = acc;
  }
}

}  // namespace tsem
