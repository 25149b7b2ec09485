// Internal definitions shared by the C ABI translation units: device
// buffers, the context and space handles, and the error-mapping macros.
#pragma once
#include "../../include/tsem.h"

#include <hip/hip_runtime.h>

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "kernels.hpp"
#include "pmg.hpp"
#include "slab.hpp"
#include "space.hpp"


#define HIPCHK(x)                                                                      \
  do {                                                                                 \
    hipError_t e_ = (x);                                                               \
    if (e_ != hipSuccess) throw std::runtime_error(std::string("HIP: ") + hipGetErrorString(e_) + " at " #x); \
  } while (0)

namespace tsem {

// last error message of the calling thread (tsem_last_error)
extern thread_local std::string g_err;

// looks up the device's copy of tsem_apply_noalloc's scratch array (the
// device must be current; once per device)
void scratch_init(int device);

struct DevBuf {
  void* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { if (p) (void)hipFree(p); }
  void alloc(size_t bytes) {
    if (p && n >= bytes) return;
    if (p) { (void)hipFree(p); p = nullptr; }
    if (bytes == 0) bytes = 8;
    if (hipMalloc(&p, bytes) != hipSuccess) { p = nullptr; throw std::bad_alloc(); }
    n = bytes;
  }
  template <typename T> T* as() const { return reinterpret_cast<T*>(p); }
};

template <typename T>
inline void upload(DevBuf& b, const std::vector<T>& v, hipStream_t st) {
  b.alloc(v.size() * sizeof(T));
  if (!v.empty()) HIPCHK(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
}

}  // namespace tsem

using tsem::DevBuf;
using tsem::DevSpace;
using tsem::g_err;
using tsem::Pmg;
using tsem::Space;
using tsem::upload;

// A context stays alive until its last space is destroyed, whatever order
// the caller releases them in.
struct tsem_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int nspaces = 0;
  bool closing = false;
  int cls = 0;  // solve class: 0 normal, 1 foreground, 2 background (tsem_ctx_set_class)
};

struct tsem_space {
  tsem_ctx* ctx = nullptr;
  Space h;
  int qorder = 0, nq1 = 0, nq2 = 0, nq3 = 0;
  std::vector<double> t1, w1, p2, w2, p3, w3;  // BC/RHS quadrature
  std::vector<double> diag;                     // host copy
  std::vector<uint8_t> cmask;
  DevSpace d;
  // device storage
  DevBuf ftc_n, ftc_key, ftc_w, ftc_w32;
  DevBuf desc, slot, ft_ptr, ft_key, ft_w, ghat, ghat32, sixv, sixv32, dT, dN, dT32, dN32, mT, mN, mT32, mN32;
  DevBuf sc_ptr, sc_code, sc_w, sc_w32, diag_d, invdiag, invdiag32, cmask_d, xe, pts3, wphi, Dp64, Dp32, Mm64, Mm32;
  DevBuf yl64, yl32, bl, partial, scal, xt64, xt32, gmask;
  DevBuf rhs_part;  // point-range partial sums of the load-vector GEMM
  // quadrature-point work: basis at the points, weights, potential, scratch
  DevBuf qphi, qphiT, qdphi, qw, vpot, qxl, qpart, qout, grad2;
  // boundary-condition quadrature and modal bases (bc.hip), offsets per table
  DevBuf bc_tab, bc_offd;
  std::vector<long> bc_off;
  DevBuf Dh1hi, Dh1lo, Dh2hi, Dh2lo;  // fp16-split FP32 operator factors
  tsem::Slab slab;                                      // entity-slab layout (host tables)
  // point location: element bounding boxes bucketed on a uniform grid
  // (tsem_point_values; dropped when the coordinates change)
  struct Locator {
    bool valid = false;
    int n[3] = {0, 0, 0};
    double lo[3] = {0, 0, 0}, h[3] = {1, 1, 1};
    std::vector<int32_t> ptr, elems;  // CSR over cells
  } loc;
  DevBuf sl_meta, sl_chunk, sl_eoff, sl_perm, sl_tk, sl_tw, sl_tw32;
  // solver vectors
  DevBuf v64[6], v32[6];
  // instantiated graph of one chunk of PCG iterations per precision, valid
  // while every pointer and parameter it captured is unchanged (key)
  struct PcgGraph {
    hipGraphExec_t exec = nullptr;
    std::vector<const void*> key;
    std::vector<const void*> failed;  // key whose capture failed: run it eagerly
    void reset() {
      if (exec) (void)hipGraphExecDestroy(exec);
      exec = nullptr;
      key.clear();
    }
    ~PcgGraph() { reset(); }
  };
  // one cached graph per chunk length class (the full check_every chunk and
  // the shorter power-of-two chunks of the last iterations) and precision
  static constexpr int GRAPH_SLOTS = 8;
  struct PcgGraphs {
    PcgGraph slot[GRAPH_SLOTS];
    void reset() {
      for (auto& g : slot) g.reset();
    }
  } graph64, graph32;
  std::unique_ptr<Pmg> pmg;
  // bumped whenever the operator state a captured PCG chunk depends on
  // changes (p-multigrid levels, eliminations, geometry): part of the graph
  // key, so no graph outlives the buffers and coefficients it captured
  uint64_t state_gen = 0;
  // drop the p-multigrid levels and every graph that captured them
  void drop_pmg() {
    pmg.reset();
    graph64.reset();
    graph32.reset();
    ++state_gen;
  }
  bool rhs_open = false;
  hipStream_t st() const {
    if (!ctx) throw std::logic_error("space was created without a device context");
    return ctx->stream;
  }
};

// upload the elimination mask and the per-element geometry bits derived from it
void upload_mask(tsem_space* sp, hipStream_t st);

#define API_BEGIN try {
#define API_END                                         \
  }                                                     \
  catch (const std::bad_alloc&) {                       \
    g_err = "allocation failed";                        \
    return TSEM_ENOMEM;                                 \
  }                                                     \
  catch (const std::invalid_argument& e) {              \
    g_err = e.what();                                   \
    return TSEM_EINVAL;                                 \
  }                                                     \
  catch (const std::logic_error& e) {                   \
    g_err = e.what();                                   \
    return TSEM_ESTATE;                                 \
  }                                                     \
  catch (const std::exception& e) {                     \
    g_err = e.what();                                   \
    return TSEM_EDEVICE;                                \
  }                                                     \
  return TSEM_OK;

