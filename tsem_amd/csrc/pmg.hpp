// p-multigrid preconditioner on the device (TSEMLib init_PMG / ite_PMG,
// src/TSEMSolver.cpp:35-136, 1119-1184).
//
// Levels of polynomial order p, p / 2, 1 by default (the reference uses p,
// p-1, ..., 1, n_project = p - 1; TSEM_PMG_ORDERS selects other coarse
// orders, TSEM_PMG_COARSE the coarsest level's smoothing degree).  The
// level-l operator is the principal submatrix of the
// eliminated operator on the DOFs of degree <= p_l (build_M_exactPartial,
// src/TSEMSolver.cpp:572-655).  The basis is hierarchical and every
// geometry lists its DOFs by degree, so those DOFs are a prefix of each
// geometry's DOFs: level l is exactly the order-p_l space on the same mesh,
// laid out as a space of that order would be, and its operator is that
// space's operator with the restricted elimination.  Each coarse level is
// therefore held as a space of its own (compact vectors of N_l DOFs, element
// kernels at K_l modes, the same matrix-free kernels as the fine level).
// Transfer is injection and zero extension through an index map per level
// (build_restrict_identity / build_interpolate_identity,
// src/TSEMSolver.cpp:206-245, 344-382).  The reference's sequential
// Gauss-Seidel sweeps become Chebyshev smoothing on the Jacobi-scaled level
// operator: parallel, and the V-cycle stays symmetric.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "kernels.hpp"

struct tsem_space;

namespace tsem {

class Pmg {
 public:
  // builds the coarse levels of `fine` (its current elimination and
  // geometry; rebuilt when either changes)
  Pmg(tsem_space* fine, hipStream_t st);
  ~Pmg();
  Pmg(const Pmg&) = delete;
  Pmg& operator=(const Pmg&) = delete;
  // z = V(r): one V-cycle from a zero initial guess; `smooth` is the
  // Chebyshev degree of the pre- and post-smoothing on every level but the
  // coarsest (which gets coarse_deg).  ncol > 1: V-cycles of ncol columns
  // (column c of r and z at c N_0) with every level's operator, smoother
  // and transfer one launch over the columns; FP32 only, and only when
  // block_supported().  Each column gets the same arithmetic as alone.
  template <typename T>
  void vcycle(const T* r, T* z, int smooth, hipStream_t st, int ncol = 1);
  // the same V-cycle captured once as a graph and replayed while r, z and
  // the parameters stay the same (eager when graphs or kernel timing are
  // on, or when the capture failed); for callers that run many V-cycles
  // outside a captured solver chunk
  template <typename T>
  void vcycle_replay(const T* r, T* z, int smooth, hipStream_t st, int ncol = 1);
  // every level's FP32 operator has a block form (fp16-split path)
  bool block_supported() const;
  int levels() const { return (int)order.size(); }
  std::vector<int> order;        // polynomial order per level (level 0 = the fine space)
  std::vector<double> lmax;      // spectral radius estimate of D^-1 A_l
  std::vector<long> ndof_level;  // N_l
  int coarse_deg = 20;
  // the coarsest level's Chebyshev interval [hi / ratio, hi]: at degree 20,
  // ratio 100 takes the 32^3, p = 6 solve from 101 to 98 iterations (mixed
  // 170 -> 166 ms, FP64 344 -> 339 ms) and leaves the eigen solves as they
  // were; 300 and degrees 12-28 were no better
  double coarse_ratio = 100.0;

 private:
  tsem_space* fine;
  std::vector<tsem_space*> sub;      // levels 1 .. L-1
  std::vector<int32_t*> map;         // level l >= 1: index in level l-1 of each level-l DOF
  std::vector<void*> vx, vb, vr, vd;  // per level: iterate, right-hand side, A x, Chebyshev step
  // the order-1 coarsest level assembled (CSR with the elimination applied:
  // eliminated rows and columns hold only the diagonal): about 15 entries
  // per row, where the matrix-free operator would run over every element
  int csr_level = -1;
  int32_t* csr_rs = nullptr;
  int32_t* csr_col = nullptr;
  double* csr_v64 = nullptr;
  float* csr_v32 = nullptr;
  void assemble_level(int l, hipStream_t st);
  const DevSpace& dev(int l) const;
  void release();
  hipGraphExec_t vexec = nullptr;
  std::vector<const void*> vkey, vfailed;
  // column blocks of the level vectors and element results (FP32, bcap
  // columns; grown by block_reserve, never inside a capture)
  int bcap = 0;
  std::vector<void*> bx, bb, br, bd, byl;
  void block_reserve(int ncol);
  long yl_cols(int l) const;  // FP32 element-result entries of one column
  template <typename T>
  void apply_level(int l, const T* x, T* y, hipStream_t st, int ncol = 1, const ChebFuse& cf = ChebFuse());
  template <typename T>
  void smooth_level(int l, T* x, const T* b, int degree, bool zero_init, hipStream_t st, int ncol = 1);
};

}  // namespace tsem
