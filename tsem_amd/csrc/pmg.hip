// p-multigrid V-cycle with Chebyshev smoothing (see pmg.hpp).
#include "pmg.hpp"

#include <cmath>
#include <random>
#include <stdexcept>

namespace tsem {

namespace {
constexpr int BLK = 256;
#define HIPCHK(x)                                                                                       \
  do {                                                                                                  \
    hipError_t e_ = (x);                                                                                \
    if (e_ != hipSuccess) throw std::runtime_error(std::string("HIP: ") + hipGetErrorString(e_)); \
  } while (0)

inline int nblk(int n) { return (n + BLK - 1) / BLK; }

// out = in where degree <= p, else 0
template <typename T>
__global__ void k_restrict(int n, const uint8_t* __restrict__ sd, int p, const T* __restrict__ in, T* __restrict__ out) {
  int i = blockIdx.x * BLK + threadIdx.x;
  if (i < n) out[i] = (sd[i] <= p) ? in[i] : T(0);
}

// Chebyshev step: d = c1 d + c2 Dinv (b - Ax) ; x += d
template <typename T>
__global__ void k_cheb(int n, const T* __restrict__ b, const T* __restrict__ Ax, const T* __restrict__ inv,
                       T c1, T c2, T* __restrict__ dv, T* __restrict__ x) {
  int i = blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  T r = b[i] - Ax[i];
  // first step (c1 = 0) must not read dv: it may hold stale data of the other precision
  T dn = c2 * inv[i] * r;
  if (c1 != T(0)) dn += c1 * dv[i];
  dv[i] = dn;
  x[i] += dn;
}

// x += y
template <typename T>
__global__ void k_add(int n, const T* __restrict__ y, T* __restrict__ x) {
  int i = blockIdx.x * BLK + threadIdx.x;
  if (i < n) x[i] += y[i];
}

__global__ void k_masked_inv(int n, const uint8_t* __restrict__ sd, int p, const double* __restrict__ invd,
                             double* __restrict__ o64, float* __restrict__ o32) {
  int i = blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  double v = (sd[i] <= p) ? invd[i] : 0.0;
  o64[i] = v;
  o32[i] = (float)v;
}

__global__ void k_scale_inv(int n, const double* __restrict__ w, const double* __restrict__ inv, double s,
                            double* __restrict__ v) {
  int i = blockIdx.x * BLK + threadIdx.x;
  if (i < n) v[i] = s * inv[i] * w[i];
}

__global__ void __launch_bounds__(BLK) k_dot2(int n, const double* __restrict__ a, const double* __restrict__ b,
                                              const double* __restrict__ inv, double* __restrict__ part) {
  __shared__ double red[2][BLK];
  int i = blockIdx.x * BLK + threadIdx.x;
  double s0 = 0, s1 = 0;
  if (i < n) { s0 = a[i] * b[i]; s1 = a[i] * a[i] / (inv[i] != 0 ? inv[i] : 1.0) * (inv[i] != 0); }
  red[0][threadIdx.x] = s0; red[1][threadIdx.x] = s1;
  __syncthreads();
  for (int s = BLK / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) { red[0][threadIdx.x] += red[0][threadIdx.x + s]; red[1][threadIdx.x] += red[1][threadIdx.x + s]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) { part[blockIdx.x] = red[0][0]; part[gridDim.x + blockIdx.x] = red[1][0]; }
}

}  // namespace

Pmg::Pmg(const Space& S, const DevSpace& dd, const std::vector<uint8_t>& cmask, hipStream_t st) : d(dd), n(dd.ndof) {
  (void)cmask;
  int p = S.M;
  for (int o = p; o >= 1; --o) order.push_back(o);
  int L = (int)order.size();
  std::vector<uint8_t> sd(n);
  for (int i = 0; i < n; ++i) sd[i] = (uint8_t)S.sum_dof[i];
  HIPCHK(hipMalloc(&sumdof, n));
  HIPCHK(hipMemcpyAsync(sumdof, sd.data(), n, hipMemcpyHostToDevice, st));
  inv64.resize(L); inv32.resize(L); work.resize(4 * L);
  ndof_level.resize(L);
  for (int l = 0; l < L; ++l) {
    long cnt = 0;
    for (int i = 0; i < n; ++i) cnt += sd[i] <= order[l];
    ndof_level[l] = cnt;
    HIPCHK(hipMalloc(&inv64[l], (size_t)n * 8));
    HIPCHK(hipMalloc(&inv32[l], (size_t)n * 4));
    hipLaunchKernelGGL(k_masked_inv, dim3(nblk(n)), dim3(BLK), 0, st, n, (const uint8_t*)sumdof, order[l], d.invdiag,
                       (double*)inv64[l], (float*)inv32[l]);
    for (int k = 0; k < 4; ++k) HIPCHK(hipMalloc(&work[4 * l + k], (size_t)n * 8));
  }
  HIPCHK(hipMalloc(&yl, (size_t)d.ntile * TILE * d.K * 8));
  // spectral radius of D^-1 A_l by power iteration (FP64)
  lmax.assign(L, 2.0);
  std::mt19937_64 rng(12345);
  std::uniform_real_distribution<double> U(-1, 1);
  std::vector<double> v0(n);
  for (int i = 0; i < n; ++i) v0[i] = U(rng);
  double* part;
  HIPCHK(hipMalloc(&part, (size_t)2 * nblk(n) * 8));
  double* fin;
  HIPCHK(hipMalloc(&fin, 16));
  for (int l = 0; l < L; ++l) {
    double* v = (double*)work[4 * l];
    double* w = (double*)work[4 * l + 1];
    HIPCHK(hipMemcpyAsync(w, v0.data(), (size_t)n * 8, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_restrict<double>, dim3(nblk(n)), dim3(BLK), 0, st, n, (const uint8_t*)sumdof, order[l],
                       (const double*)w, v);
    double lam = 1.0;
    for (int it = 0; it < 25; ++it) {
      apply_level<double>(l, v, w, st);
      // w = Dinv A v ; estimate lambda = (v, A v) / (v, D v)  (Rayleigh quotient in the D inner product)
      hipLaunchKernelGGL(k_dot2, dim3(nblk(n)), dim3(BLK), 0, st, n, (const double*)v, (const double*)w,
                         (const double*)inv64[l], part);
      launch_reduce_final(part, nblk(n), 2, fin, st);
      double h[2];
      HIPCHK(hipMemcpyAsync(h, fin, 16, hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      lam = h[0] / h[1];
      // v <- Dinv A v, normalized in the D norm
      hipLaunchKernelGGL(k_scale_inv, dim3(nblk(n)), dim3(BLK), 0, st, n, (const double*)w, (const double*)inv64[l],
                         1.0 / std::sqrt(h[1]) / (lam > 0 ? lam : 1.0), v);
    }
    lmax[l] = lam;
  }
  HIPCHK(hipFree(part));
  HIPCHK(hipFree(fin));
  HIPCHK(hipStreamSynchronize(st));
}

Pmg::~Pmg() {
  for (void* p : inv64) (void)hipFree(p);
  for (void* p : inv32) (void)hipFree(p);
  for (void* p : work) (void)hipFree(p);
  if (sumdof) (void)hipFree(sumdof);
  if (yl) (void)hipFree(yl);
}

template <typename T>
void Pmg::apply_level(int l, const T* x, T* y, hipStream_t st) {
  OpCoef op;
  op.constrained = 1;
  op.Kl = ModeIndex::count_upto(order[l]);
  op.Kdl = ModeIndex::count_upto(order[l] - 1);
  launch_apply<T>(d, op, x, y, (T*)yl, nullptr, st);
}

// Chebyshev iteration of given degree on D^-1 A_l x = D^-1 b over the interval
// [lmax/10, 1.1 lmax]
template <typename T>
void Pmg::smooth_level(int l, T* x, const T* b, int degree, bool zero_init, hipStream_t st) {
  T* Ax = (T*)work[4 * l + 2];
  T* dv = (T*)work[4 * l + 3];
  const T* inv = (sizeof(T) == 8) ? (const T*)inv64[l] : (const T*)inv32[l];
  double hi = 1.1 * lmax[l], lo = lmax[l] / 10.0;
  double theta = (hi + lo) / 2, delta = (hi - lo) / 2, sigma = theta / delta;
  double rho_old = 1.0 / sigma;
  if (zero_init) launch_zero<T>(n, Ax, st);
  else apply_level<T>(l, x, Ax, st);
  // first step: d = Dinv r / theta
  hipLaunchKernelGGL(k_cheb<T>, dim3(nblk(n)), dim3(BLK), 0, st, n, b, (const T*)Ax, inv, T(0), T(1.0 / theta), dv, x);
  for (int k = 1; k < degree; ++k) {
    apply_level<T>(l, x, Ax, st);
    double rho = 1.0 / (2 * sigma - rho_old);
    hipLaunchKernelGGL(k_cheb<T>, dim3(nblk(n)), dim3(BLK), 0, st, n, b, (const T*)Ax, inv, T(rho * rho_old),
                       T(2 * rho / delta), dv, x);
    rho_old = rho;
  }
}

template <typename T>
void Pmg::vcycle(const T* r, T* z, int smooth, hipStream_t st) {
  int L = levels();
  if (smooth < 1) smooth = 1;
  // level 0 uses z as its iterate and r as its right-hand side
  std::vector<T*> X(L), B(L);
  X[0] = z;
  B[0] = const_cast<T*>(r);
  for (int l = 1; l < L; ++l) { X[l] = (T*)work[4 * l]; B[l] = (T*)work[4 * l + 1]; }
  launch_zero<T>(n, z, st);
  for (int l = 0; l + 1 < L; ++l) {
    smooth_level<T>(l, X[l], B[l], smooth, true, st);
    T* Ax = (T*)work[4 * l + 2];
    apply_level<T>(l, X[l], Ax, st);
    // B[l+1] = restrict(B[l] - A X[l])
    launch_axpby<T>(n, T(1), B[l], T(-1), Ax, st);
    hipLaunchKernelGGL(k_restrict<T>, dim3(nblk(n)), dim3(BLK), 0, st, n, (const uint8_t*)sumdof, order[l + 1],
                       (const T*)Ax, B[l + 1]);
    launch_zero<T>(n, X[l + 1], st);
  }
  smooth_level<T>(L - 1, X[L - 1], B[L - 1], 2 * smooth + 2, true, st);
  for (int l = L - 2; l >= 0; --l) {
    hipLaunchKernelGGL(k_add<T>, dim3(nblk(n)), dim3(BLK), 0, st, n, (const T*)X[l + 1], X[l]);
    smooth_level<T>(l, X[l], B[l], smooth, false, st);
  }
}

template void Pmg::vcycle<double>(const double*, double*, int, hipStream_t);
template void Pmg::vcycle<float>(const float*, float*, int, hipStream_t);

}  // namespace tsem
