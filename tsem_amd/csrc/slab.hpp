// Entity-slab layout of element results (host construction).
//
// The element kernels can write their results already transposed through the
// gather map, one row per (geometry, incident element): a geometry g of
// dimension d with n_d DOFs starting at global DOF s and m incident elements
// owns the slab range [P[s], P[s] + m n_d), P the exclusive prefix sum over
// global DOFs of their element counts; the element of rank k (elements
// counted in element order) writes its n_d values at P[s] + k n_d.  The value
// written for global DOF s + i is the element's whole contribution to that
// DOF: its own mode for vertices and interiors, the edge mode with the edge
// sign plus the face modes that the face re-expression sends to the edge, or
// the face modes that reach the face DOF.  The scatter then adds, per DOF, m
// slab entries spaced n_d apart, with no codes or weights.  The slab has
// nt K entries, as many as the element-major results.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "space.hpp"

namespace tsem {

constexpr int SLAB_ROWS = 256;  // DOFs per scatter chunk
// the chunk table is padded with empty chunks to a multiple of this many, so
// that a scatter step reads the bounds of a whole group without a test
constexpr int SLAB_CHUNK_GROUP = 8;
constexpr int SLAB_JMAX = 8;    // most terms of a transposed face row

// rows of the transposed face tables: face DOF rows (v, i), then edge DOF
// rows (v, ie, l)
inline int slab_rows(int n1, int n2) { return 6 * n2 + 18 * n1 > 0 ? 6 * n2 + 18 * n1 : 1; }

struct Slab {
  bool ok = false;
  std::string why;  // reason when !ok
  int64_t total = 0;
  int n1 = 0, n2 = 0, n3 = 0, JT = 1, nrow = 1;
  int chunk_max = 0;            // largest slab span of a chunk
  std::vector<int32_t> meta;    // per DOF: offset in its chunk's span | m << 20 | dim << 30
  std::vector<int32_t> chunk;   // per chunk of SLAB_ROWS DOFs: span start, span length (then
                                // empty chunks up to a multiple of SLAB_CHUNK_GROUP)
  // per element (17 ints): 15 entity offsets (v0..3, e0..5, f0..3,
  // interior), orientation words ow and ow2 (slab_edge)
  std::vector<int32_t> eoff;
  std::vector<int32_t> perm;    // [K]: slab row of local mode i
  // transposed face re-expression, rows padded to JT terms: face DOF rows
  // (v, i) first (6 n2), then edge DOF rows (v, ie, l) (18 n1); a term is a
  // face mode d of the element face (tk, the row's term count << 8 added to
  // its first entry) and a weight
  std::vector<uint16_t> tk;
  std::vector<double> tw;
  // face variants whose re-expression is diagonal: each face DOF i is +-
  // the element's face mode i, no edge terms (bit v of diag; variant 0, the
  // identity, always); sgn[v] bit i set where that sign is minus
  int diag = 0;
  uint32_t sgn[6] = {0, 0, 0, 0, 0, 0};
};

// orientation word of an element: per face f, variant (3 bits) at 6 f and the
// reversed bits of its edges ie (3 bits) at 6 f + 3; edge flips at bit 24
inline int slab_face_var(int32_t w, int f) { return (w >> (6 * f)) & 7; }

void build_slab(const Space& S, Slab& out);

// y = the scatter of element-major results yl (nt x K) through the slab, on
// the host: every element's slab entries (the sums the element kernel's
// epilogue forms, term by term), then per DOF its entries in element order
void slab_apply_host(const Space& S, const Slab& L, const double* yl, double* y);

}  // namespace tsem
