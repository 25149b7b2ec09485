// p-multigrid preconditioner on device: levels of polynomial order
// p, p-1, ..., 1 (TSEMLib init_PMG, src/TSEMSolver.cpp:35-136).  The level-l
// operator is the principal submatrix of the eliminated operator on the DOFs
// of degree <= p_l (build_M_exactPartial, src/TSEMSolver.cpp:572-655), applied
// matrix-free with the leading modes of the same reference factors; transfer
// is injection / zero extension (src/TSEMSolver.cpp:206-245, 344-382).  The
// sequential Gauss-Seidel sweeps of ite_PMG (src/TSEMSolver.cpp:1119-1184)
// become Chebyshev smoothing on the Jacobi-scaled level operator, which is
// parallel and keeps the V-cycle symmetric.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "kernels.hpp"
#include "space.hpp"

namespace tsem {

class Pmg {
 public:
  Pmg(const Space& S, const DevSpace& d, const std::vector<uint8_t>& cmask, hipStream_t st);
  ~Pmg();
  // z = V(r): one V-cycle from a zero initial guess
  template <typename T>
  void vcycle(const T* r, T* z, int smooth, hipStream_t st);
  int levels() const { return (int)order.size(); }
  std::vector<int> order;       // polynomial order per level
  std::vector<double> lmax;     // spectral radius estimate of D^-1 A_l
  std::vector<long> ndof_level;

 private:
  const DevSpace& d;
  int n;
  void* sumdof = nullptr;        // uint8 per DOF
  std::vector<void*> inv64, inv32;  // masked inverse diagonal per level
  std::vector<void*> work;       // per level x, b, r, dvec (double and float views share bytes)
  void* yl = nullptr;            // element-local scratch (double sized)
  template <typename T>
  void apply_level(int l, const T* x, T* y, hipStream_t st);
  template <typename T>
  void smooth_level(int l, T* x, const T* b, int degree, bool zero_init, hipStream_t st);
};

}  // namespace tsem
