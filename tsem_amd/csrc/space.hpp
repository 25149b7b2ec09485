// Host-side space setup: mesh topology, first-visit orientation, global DOF
// layout and the compact element descriptors the device gather/scatter use.
#pragma once
#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "refel.hpp"

namespace tsem {

// Face re-orientation table (TSEMLib include/quad_info/Info_Interpolate_More,
// read by src/BuildSpace.cpp:1703-1771): local face DOF d under variant v is
// sum_ie sum (l, w) w * edge_ie DOF l  +  sum (j, w) w * face DOF j.
struct FaceTable {
  int mmax = 0, nface = 0;
  struct Term { int ie, idx; double w; };  // ie = 0..2 edge terms, ie = 3 face terms
  std::vector<std::vector<Term>> terms;    // [v * nface + d]
  void load(const std::string& path);
};

struct Topology {
  int nv = 0, ne = 0, nf = 0, nt = 0;
  std::vector<int32_t> edge_vtx, face_vtx, face_edge, tet_face, tet_edge;
};

// Canonical topology of a tetrahedral mesh given by its element vertex lists:
// face k of a tet is opposite its vertex k, its vertices in the order
// (1,2,3), (0,2,3), (0,1,3), (0,1,2); edges E01,E02,E03,E12,E13,E23; faces
// and edges are numbered by first appearance scanning the tets in order, and
// keep the vertex order of that first appearance; face_edge[f][j] is the
// edge opposite face vertex j.
void build_topology(const int32_t* tet, int nt, int nv, Topology& topo);

// P2 DOF order of a mesh (weights that order the geometries): scan tets in
// order, number their 4 vertices then their 6 edges on first appearance.
void canonical_femdof(const Topology& topo, const int32_t* tet, std::vector<int32_t>& w);

// Per-element descriptor consumed by the device gather: global DOF starts
// of the element's geometries plus orientation codes.
struct ElemDesc {
  int32_t v[4];   // vertex DOF
  int32_t e[6];   // edge DOF start; bit 31 set when the edge runs reversed
  int32_t f[4];   // face DOF start
  int32_t fo[4];  // face code: variant (3 bits) | local edge of face edge ie (3 bits each) << 3+3ie | reversed(ie) << 12+ie
  int32_t t;      // interior DOF start
  int32_t pad[3];
};

struct Space {
  int M = 0, K = 0;
  Topology topo;
  RefElement ref;
  FaceTable table;
  int ndg[4] = {0, 0, 0, 0};
  int64_t ndof = 0;
  std::vector<double> xyz;
  std::vector<int32_t> tet;
  std::vector<int32_t> femdof;
  // first-visit orientation (src/BuildSpace.cpp:435-655)
  std::vector<std::array<int32_t, 2>> edge_node;   // global edge start/end vertex
  std::vector<std::array<int32_t, 3>> face_node;   // global face vertex order
  std::vector<std::array<int32_t, 3>> face_edges;  // global face edge ie
  std::vector<uint8_t> face_rev;                   // bit ie: edge ie of face runs against the face order
  std::vector<uint8_t> tet_eflip;                  // bit k: element edge k reversed
  std::vector<uint8_t> tet_fvar;                   // 4 x 3 bits? stored per face below
  std::vector<std::array<uint8_t, 4>> tet_fv;      // variant of each element face (inverse projection)
  std::vector<double> vol;
  // DOF layout
  std::vector<int64_t> dof_start;  // per geometry (vertices, edges, faces, tets)
  std::vector<ElemDesc> desc;
  // slot tables for the element modes
  std::vector<int32_t> slot_kind, slot_geo, slot_sub;  // per local mode
  std::vector<int32_t> sum_dof;                         // per global DOF
  // geometric factors: per element Ghat (11,22,33,12,13,23) and 6V
  std::vector<double> ghat, sixv;
  std::vector<double> pairw;  // the 6 pair weights per element (reference order)
  // transposed gather map (scatter): CSR by global DOF
  std::vector<int64_t> sc_ptr;
  std::vector<int32_t> sc_code;  // slot index | bit30 negative | bit31 weighted
  std::vector<double> sc_w;
  // boundary marks per vertex/edge/face and the Dirichlet set
  std::vector<int32_t> mark[3];

  void build(int M, int nv, const double* xyz, int nt, const int32_t* tet, const int32_t* femdof_w,
             const std::string& table_path);
  int64_t geo_dof(int dim, int i) const;
  void set_geometry(const double* xyz);
  // ragged gather map of element e, local mode i: appends (global, weight)
  void gather_terms(int e, int i, std::vector<std::pair<int64_t, double>>& out) const;
  // assembled operator diagonal of  stiff * S + mass_coef * M
  void diagonal(double stiff, double mass_coef, std::vector<double>& diag) const;
  // element matrix sum_ab Ghat_ab grad2_ab (+ mass term)
  void element_matrix(int e, double stiff, double mass_coef, std::vector<double>& Ke) const;
};

}  // namespace tsem
