// tsem_apply_noalloc: the matrix-free operator of tsem_apply with its
// element-local scratch (gathered element tiles, element results or entity
// slab) in a file-scope device array instead of the space's own buffers, for
// callers that must not allocate device memory while they run.
//
// The array is a __device__ symbol: in host code its name is a stand-in, not
// a device address, so its address on each device is looked up once
// (hipGetSymbolAddress, with that device current) when a context for the
// device is created, and only that address is handed to the kernels.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <stdexcept>
#include <string>

#include "handle.hpp"

namespace tsem {

constexpr size_t SCRATCH_BYTES = size_t(64) << 20;  // per device, held until the process ends
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[MAX_DEVICES] = {nullptr};

inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }
}  // namespace

// Looks up this device's copy of the scratch array; the device must be
// current.  Called by tsem_ctx_create for its device, and by the entry point
// for a device no context set up.
void scratch_init(int device) {
  if (device < 0 || device >= MAX_DEVICES) throw std::invalid_argument("device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return;
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    throw std::runtime_error(std::string("HIP: scratch symbol lookup failed: ") + hipGetErrorString(e));
  g_scratch_ptr[device] = p;
}

static void* scratch_of(int device) {
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (device >= 0 && device < MAX_DEVICES && g_scratch_ptr[device]) return g_scratch_ptr[device];
  }
  HIPCHK(hipSetDevice(device));
  scratch_init(device);
  return g_scratch_ptr[device];
}

// Scratch layout of one apply: element results (or entity slab) first, then
// the gathered tiles of the element kernel the operator will run.  Returns the
// bytes needed; with base set, points the kernels' tile pointers of d into it.
template <typename T>
static size_t scratch_layout(DevSpace& d, const OpCoef& op, char* base, T** yl) {
  const size_t tiles = (size_t)d.ntile * TILE;
  size_t off = align256(tiles * d.K * sizeof(T));  // results: at most one value per (element, mode)
  if (base) *yl = reinterpret_cast<T*>(base);
  if (sizeof(T) == 4 && split_path(d, op)) {
    // the fp16-split kernel gathers into LDS: results only
  } else if (op.mass == 0.0 && (apply_path() == 0 || apply_path() == 2) && mfma_supported(d)) {
    const size_t t = tiles * d.Kp * sizeof(T);
    if (base) {
      if (sizeof(T) == 8) d.xt64 = reinterpret_cast<double*>(base + off);
      else d.xt32 = reinterpret_cast<float*>(base + off);
    }
    off += align256(t);
  }
  return off;
}

size_t scratch_bytes() { return SCRATCH_BYTES; }

}  // namespace tsem

using namespace tsem;

extern "C" int tsem_apply_noalloc(tsem_space* sp, double stiff, double mass, int constrained, int prec, const void* x,
                                  void* y) {
  API_BEGIN
  if (!sp || !x || !y) throw std::invalid_argument("null argument");
  if (prec != 64 && prec != 32) throw std::invalid_argument("prec must be 64 or 32");
  hipStream_t st = sp->st();
  char* base = static_cast<char*>(tsem::scratch_of(sp->ctx->device));
  OpCoef op;
  op.stiff = stiff; op.mass = mass; op.constrained = constrained;
  DevSpace d = sp->d;  // the space's device data with the tile pointers moved into the scratch array
  size_t need;
  if (prec == 64) {
    double* yl = nullptr;
    need = scratch_layout<double>(d, op, nullptr, nullptr);
    if (need > tsem::scratch_bytes())
      throw std::invalid_argument("space too large for tsem_apply_noalloc: needs " + std::to_string(need) +
                                  " scratch bytes, the limit is " + std::to_string(tsem::scratch_bytes()));
    scratch_layout<double>(d, op, base, &yl);
    launch_apply<double>(d, op, (const double*)x, (double*)y, yl, nullptr, st);
  } else {
    float* yl = nullptr;
    need = scratch_layout<float>(d, op, nullptr, nullptr);
    if (need > tsem::scratch_bytes())
      throw std::invalid_argument("space too large for tsem_apply_noalloc: needs " + std::to_string(need) +
                                  " scratch bytes, the limit is " + std::to_string(tsem::scratch_bytes()));
    scratch_layout<float>(d, op, base, &yl);
    launch_apply<float>(d, op, (const float*)x, (float*)y, yl, nullptr, st);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(st));
  API_END
}
