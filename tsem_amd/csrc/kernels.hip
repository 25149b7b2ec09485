// HIP kernels of the TSEM operator path for gfx950.
//
//  gather   : element-local coefficients from global DOFs through the compact
//             element descriptors and the face re-orientation table
//             (TSEMLib transform_*_global2local, src/BuildSpace.cpp:1021-1101);
//  apply    : y_e = sum_ab Ghat_ab D_a^T D_b x_e  (+ mass term), the element
//             operator sum_k w_k S_k of src/Discretize.cpp:1352-1393 written
//             through an orthonormal factorization of the reference stiffness;
//  scatter  : owner-computes sum over the transposed gather map, fixed order,
//             no atomics (bit-identical run to run);
//  CG/vector: fused PCG updates with deterministic two-stage reductions.
#include <hip/hip_runtime.h>

#include <stdexcept>

#include <cstdlib>
#include <utility>
#include <vector>

#include "kernels.hpp"
#include "slab.hpp"

namespace tsem {

constexpr int BLK = 256;

// ---------------------------------------------------------------- gather ----
template <typename T>
__device__ inline T ldx(const T* __restrict__ x, const uint8_t* __restrict__ cmask, int g, int constrained) {
  if (constrained && cmask[g]) return T(0);
  return x[g];
}

// one local coefficient; FP64 sums are kept free of fused multiply-adds so the
// result is bit-identical to a plain sequential evaluation
template <typename T>
__device__ inline T gather_slot(const DevSpace& S, int e, int i, const T* __restrict__ x, int constrained) {
#pragma clang fp contract(off)
  const ElemDesc& D = S.desc[e];
  int code = S.slot[i];
  int kind = code & 3, geo = (code >> 2) & 7, sub = code >> 5;
  if (kind == 0) return ldx(x, S.cmask, D.v[geo], constrained);
  if (kind == 1) {
    int ed = D.e[geo];
    T v = ldx(x, S.cmask, (ed & 0x7fffffff) + sub, constrained);
    // reversed edge: odd-degree modes (l = sub + 2 odd) change sign
    return (ed < 0 && (sub & 1)) ? -v : v;
  }
  if (kind == 3) return ldx(x, S.cmask, D.t + sub, constrained);
  int fo = D.fo[geo];
  int var = fo & 7;
  int row = var * S.nface + sub;
  T acc = T(0);
  for (int t = S.ft_ptr[row]; t < S.ft_ptr[row + 1]; ++t) {
    int key = S.ft_key[t];
    int ie = key & 3, idx = key >> 2;
    T w = (T)S.ft_w[t];
    int g;
    if (ie < 3) {
      int le = (fo >> (3 + 3 * ie)) & 7;
      if (((fo >> (12 + ie)) & 1) && (idx & 1)) w = -w;
      g = (D.e[le] & 0x7fffffff) + idx;
    } else {
      g = D.f[geo] + idx;
    }
    T xv = ldx(x, S.cmask, g, constrained);
    acc += xv * w;
  }
  return acc;
}

__global__ void __launch_bounds__(BLK) k_gather(DevSpace S, const double* __restrict__ x, double* __restrict__ xl,
                                                int constrained) {
  long s = (long)blockIdx.x * BLK + threadIdx.x;
  long n = (long)S.nt * S.K;
  if (s >= n) return;
  int e = (int)(s / S.K), i = (int)(s % S.K);
  xl[s] = gather_slot<double>(S, e, i, x, constrained);
}

void launch_gather(const DevSpace& S, const double* x, double* xl, int constrained, hipStream_t st) {
  long n = (long)S.nt * S.K;
  hipLaunchKernelGGL(k_gather, dim3((unsigned)((n + BLK - 1) / BLK)), dim3(BLK), 0, st, S, x, xl, constrained);
}

// yl[e][i] += pot vol_e sum_q w_q V[e][q] phi_i(x_q) sum_j phi_j(x_q) x_e[j]:
// the potential mass operator M_V applied matrix-free (the assembled form is
// TSEM::build_mass_V_matrix, src/Discretize.cpp:667-843).  One block per
// element; the element coefficients and the weighted point values sit in LDS.
template <typename T>
__global__ void __launch_bounds__(BLK) k_apply_potential(DevSpace S, double pot, const T* __restrict__ x,
                                                         T* __restrict__ yl, int constrained) {
  extern __shared__ double psm[];
  double* xe = psm;         // K
  double* c = psm + S.K;    // nq
  const int e = blockIdx.x, K = S.K, nq = S.nq;
  for (int i = threadIdx.x; i < K; i += BLK) xe[i] = (double)gather_slot<T>(S, e, i, x, constrained);
  __syncthreads();
  const double scale = pot * S.sixv[e] / 6.0;
  const double* V = S.vpot + (size_t)e * nq;
  for (int q = threadIdx.x; q < nq; q += BLK) {
    double u = 0.0;
    for (int i = 0; i < K; ++i) u += S.phiT[(size_t)i * nq + q] * xe[i];
    c[q] = scale * S.wq[q] * V[q] * u;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < K; i += BLK) {
    double y = 0.0;
    for (int q = 0; q < nq; ++q) y += S.phi[(size_t)q * K + i] * c[q];
    yl[(size_t)e * K + i] += (T)y;
  }
}

// M_V on the matrix cores (v_mfma_f64_16x16x4f64), 64 elements per block and
// wave w owning elements 16 w .. 16 w + 15.  For each stage of 16 points,
//   U^T (16 x 16) = Phi (16 points x K) X^T (K x 16 elements),
//   C = U^T scaled by pot vol_e w_q V[e][q],
//   Y (16 x K) += C^T (16 x 16) Phi (16 x K).
// The first product's accumulator (value j of lane l: point l/16 + 4 j,
// element l%16) is the second product's A operand as it stands: step j of the
// second product takes point 4 j + l/16 as its k index in both operands.
// X stays in registers (value s of lane l: element l%16, mode 4 s + l/16);
// the stage's Phi rows and weighted point values sit in LDS.  Every point
// costs each element 4 K FLOPs, which bounds the kernel by the FP64 matrix
// rate (V, nq doubles per element, is the only large HBM stream).  Measured
// at p = 5 (nq = 512) on 82,944 elements: stages of 16 points beat 32 and 64
// (fewer registers, 3 waves per SIMD); gathering X through LDS, prefetching
// the next stage into registers and a 4-wave register cap were all slower.
typedef double pd4 __attribute__((ext_vector_type(4)));
constexpr int POT_PE = 64, POT_RQ = 16;

// up to 4 mode tiles the kernel is held to 128 registers (4 waves per SIMD,
// two spilled): at p = 5 the apply went 543 -> 525 us; wider tiles spill too
// much under that bound (p = 6: 1,122 -> 1,262 us)
template <typename T, int MT>
__global__ void __launch_bounds__(BLK, MT <= 4 ? 4 : 1) k_pot_gemm(DevSpace S, double pot, const T* __restrict__ x,
                                                  T* __restrict__ yl, int constrained) {
  constexpr int NW = 16 * MT, RQ = POT_RQ;
  constexpr int PP = RQ * NW / BLK, PC = POT_PE * RQ / BLK;  // stage loads per thread
  __shared__ double sP[RQ][NW + 1];      // phi_i(x_q) of the stage's points
  __shared__ double sC[POT_PE][RQ + 1];  // pot vol_e w_q V[e][q]
  __shared__ double sS[POT_PE];          // pot vol_e
  const int K = S.K, nq = S.nq, nks = (K + 3) >> 2;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, lr = lane & 15, lk = lane >> 4;
  const int eb = xcd_block(blockIdx.x, gridDim.x) * POT_PE;
  const int ne = min(POT_PE, S.nt - eb);
  const int ew = eb + 16 * wave + lr;
  if (tid < POT_PE) sS[tid] = (tid < ne) ? pot * S.sixv[eb + tid] / 6.0 : 0.0;
  double xr[4 * MT];
#pragma unroll
  for (int s = 0; s < 4 * MT; ++s) {
    const int i = 4 * s + lk;
    xr[s] = (i < K && ew < S.nt) ? (double)gather_slot<T>(S, ew, i, x, constrained) : 0.0;
  }
  pd4 acc[MT];
#pragma unroll
  for (int t = 0; t < MT; ++t) acc[t] = pd4{0, 0, 0, 0};
  __syncthreads();
  for (int q0 = 0; q0 < nq; q0 += RQ) {
    double rp[PP], rc[PC];
#pragma unroll
    for (int k = 0; k < PP; ++k) {
      const int s = tid + k * BLK, qq = s / NW, i = s - qq * NW, q = q0 + qq;
      rp[k] = (q < nq && i < K) ? S.phi[(size_t)q * K + i] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < PC; ++k) {
      const int s = tid + k * BLK, le = s / RQ, qq = s - le * RQ, q = q0 + qq;
      rc[k] = (le < ne && q < nq) ? S.vpot[(size_t)(eb + le) * nq + q] * S.wq[q] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < PP; ++k) {
      const int s = tid + k * BLK, qq = s / NW;
      sP[qq][s - qq * NW] = rp[k];
    }
#pragma unroll
    for (int k = 0; k < PC; ++k) {
      const int s = tid + k * BLK, le = s / RQ;
      sC[le][s - le * RQ] = rc[k] * sS[le];
    }
    __syncthreads();
    pd4 u = pd4{0, 0, 0, 0};
#pragma unroll
    for (int s = 0; s < 4 * MT; ++s)
      if (s < nks) u = __builtin_amdgcn_mfma_f64_16x16x4f64(sP[lr][4 * s + lk], xr[s], u, 0, 0, 0);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int qq = lk + 4 * j;
      const double a = u[j] * sC[16 * wave + lr][qq];
#pragma unroll
      for (int t = 0; t < MT; ++t)
        acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, sP[qq][16 * t + lr], acc[t], 0, 0, 0);
    }
    __syncthreads();
  }
  // value j of acc[t] in lane l: element 16 wave + l/16 + 4 j, mode 16 t + l%16
#pragma unroll
  for (int t = 0; t < MT; ++t)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int le = 16 * wave + lk + 4 * j, i = 16 * t + lr;
      if (le < ne && i < K) yl[(size_t)(eb + le) * K + i] += (T)acc[t][j];
    }
}

template <typename T, int MT>
static void pot_gemm_mt(const DevSpace& S, double pot, const T* x, T* yl, int constrained, hipStream_t st) {
  const int nb = (S.nt + POT_PE - 1) / POT_PE;
  hipLaunchKernelGGL((k_pot_gemm<T, MT>), dim3(nb), dim3(BLK), 0, st, S, pot, x, yl, constrained);
}

// M_V x added to the element results yl: the matrix-core kernel up to 128
// modes (order 7), the one-block-per-element kernel above that
template <typename T>
static void launch_potential(const DevSpace& S, double pot, const T* x, T* yl, int constrained, hipStream_t st) {
  const int mt = (S.K + 15) / 16;
  switch (mt) {
    case 1: pot_gemm_mt<T, 1>(S, pot, x, yl, constrained, st); return;
    case 2: pot_gemm_mt<T, 2>(S, pot, x, yl, constrained, st); return;
    case 3: pot_gemm_mt<T, 3>(S, pot, x, yl, constrained, st); return;
    case 4: pot_gemm_mt<T, 4>(S, pot, x, yl, constrained, st); return;
    case 5: pot_gemm_mt<T, 5>(S, pot, x, yl, constrained, st); return;
    case 6: pot_gemm_mt<T, 6>(S, pot, x, yl, constrained, st); return;
    case 7: pot_gemm_mt<T, 7>(S, pot, x, yl, constrained, st); return;
    case 8: pot_gemm_mt<T, 8>(S, pot, x, yl, constrained, st); return;
    default: break;
  }
  size_t lds = (size_t)(S.K + S.nq) * sizeof(double);
  if (lds > 64 * 1024) throw std::invalid_argument("quadrature too fine for the potential operator (qorder <= 20)");
  hipLaunchKernelGGL(k_apply_potential<T>, dim3(S.nt), dim3(BLK), lds, st, S, pot, x, yl, constrained);
}

// ----------------------------------------------------------------- apply ----
template <typename T> struct Fac;
template <> struct Fac<double> {
  static __device__ const double* dT(const DevSpace& S) { return S.dT; }
  static __device__ const double* dN(const DevSpace& S) { return S.dN; }
  static __device__ const double* mT(const DevSpace& S) { return S.mT; }
  static __device__ const double* mN(const DevSpace& S) { return S.mN; }
  static __device__ double g(const DevSpace& S, int e, int k) { return S.ghat[6 * e + k]; }
  static __device__ double v6(const DevSpace& S, int e) { return S.sixv[e]; }
};
template <> struct Fac<float> {
  static __device__ const float* dT(const DevSpace& S) { return S.dT32; }
  static __device__ const float* dN(const DevSpace& S) { return S.dN32; }
  static __device__ const float* mT(const DevSpace& S) { return S.mT32; }
  static __device__ const float* mN(const DevSpace& S) { return S.mN32; }
  static __device__ float g(const DevSpace& S, int e, int k) { return S.ghat32[6 * e + k]; }
  static __device__ float v6(const DevSpace& S, int e) { return S.sixv32[e]; }
};

// NE elements per block; LDS: X[NE][K] | U[NE][3Kd] | Wm[NE][K] (mass)
template <typename T, int NE>
__global__ void __launch_bounds__(BLK) k_apply(DevSpace S, OpCoef op, const T* __restrict__ x, T* __restrict__ yl) {
  extern __shared__ unsigned char smem_raw[];
  T* X = reinterpret_cast<T*>(smem_raw);
  const int K = S.K, Kd = S.Kd, R = 3 * Kd;
  const int Kl = op.Kl > 0 ? op.Kl : K;
  const int Kdl = op.Kdl > 0 ? op.Kdl : Kd;
  T* U = X + NE * K;
  T* W = U + NE * R;
  const int e0 = blockIdx.x * NE;
  const int tid = threadIdx.x;
  for (int s = tid; s < NE * K; s += BLK) {
    int le = s / K, i = s % K, e = e0 + le;
    X[s] = (e < S.nt && i < Kl) ? gather_slot<T>(S, e, i, x, op.constrained) : T(0);
  }
  __syncthreads();
  const T* dT = Fac<T>::dT(S);
  const T* dN = Fac<T>::dN(S);
  if (op.stiff != 0.0) {
    // U[e][a Kd + m] = sum_i D_a[m][i] X[e][i]
    for (int r = tid; r < R; r += BLK) {
      int m = r % Kd;
      T acc[NE];
#pragma unroll
      for (int le = 0; le < NE; ++le) acc[le] = T(0);
      if (m < Kdl) {
        for (int i = 0; i < Kl; ++i) {
          T d = dT[(size_t)i * R + r];
#pragma unroll
          for (int le = 0; le < NE; ++le) acc[le] += d * X[le * K + i];
        }
      }
#pragma unroll
      for (int le = 0; le < NE; ++le) U[le * R + r] = acc[le];
    }
    __syncthreads();
    // V_a = sum_b Ghat_ab U_b
    for (int s = tid; s < NE * Kd; s += BLK) {
      int le = s / Kd, m = s % Kd, e = e0 + le;
      if (e >= S.nt) continue;
      T g0 = Fac<T>::g(S, e, 0), g1 = Fac<T>::g(S, e, 1), g2 = Fac<T>::g(S, e, 2);
      T g3 = Fac<T>::g(S, e, 3), g4 = Fac<T>::g(S, e, 4), g5 = Fac<T>::g(S, e, 5);
      T c = (T)op.stiff;
      T u0 = U[le * R + m], u1 = U[le * R + Kd + m], u2 = U[le * R + 2 * Kd + m];
      U[le * R + m] = c * (g0 * u0 + g3 * u1 + g4 * u2);
      U[le * R + Kd + m] = c * (g3 * u0 + g1 * u1 + g5 * u2);
      U[le * R + 2 * Kd + m] = c * (g4 * u0 + g5 * u1 + g2 * u2);
    }
    __syncthreads();
  }
  const bool has_mass = op.mass != 0.0;
  if (has_mass) {
    const T* mT = Fac<T>::mT(S);
    for (int r = tid; r < K; r += BLK) {
      T acc[NE];
#pragma unroll
      for (int le = 0; le < NE; ++le) acc[le] = T(0);
      if (r < Kl)
        for (int i = 0; i < Kl; ++i) {
          T d = mT[(size_t)i * K + r];
#pragma unroll
          for (int le = 0; le < NE; ++le) acc[le] += d * X[le * K + i];
        }
#pragma unroll
      for (int le = 0; le < NE; ++le) {
        int e = e0 + le;
        W[le * K + r] = (e < S.nt) ? acc[le] * (T)op.mass * Fac<T>::v6(S, e) : T(0);
      }
    }
    __syncthreads();
  }
  const T* mN = Fac<T>::mN(S);
  for (int s = tid; s < NE * K; s += BLK) {
    int le = s / K, i = s % K, e = e0 + le;
    if (e >= S.nt) continue;
    T acc = T(0);
    if (i < Kl) {
      if (op.stiff != 0.0)
        for (int a = 0; a < 3; ++a)
          for (int m = 0; m < Kdl; ++m) acc += dN[(size_t)(a * Kd + m) * K + i] * U[le * R + a * Kd + m];
      if (has_mass)
        for (int m = 0; m < Kl; ++m) acc += mN[(size_t)m * K + i] * W[le * K + m];
    }
    yl[(size_t)e * K + i] = acc;
  }
}

// ---------------------------------------------------------------- scatter ----
// block reduction of one double per thread (fixed tree)
// block-wide sum in a fixed order: an xor butterfly inside each wave (every
// lane ends with the same bits: each step adds the same two values, in either
// order), then the wave sums in wave order; two barriers instead of one per
// tree level
__device__ inline double block_sum(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = 0.0;
#pragma unroll
  for (int w = 0; w < BLK / 64; ++w) r += red[w];
  __syncthreads();  // red is reused by the next call
  return r;
}

// y[g] = the row's terms of the transposed gather map summed in map order
// (or diag * x on eliminated rows).  Rows keep their natural order across
// lanes, so neighbouring lanes often read the same element-result lines; each
// lane issues the code loads of up to U terms, then their result loads, then
// adds in order.  The optional partial sums of x.y (one per block) feed the
// CG denominator.  The grid is reduce_blocks(ndof) with a fixed row-to-block
// assignment.
// the scatter's result for DOF g: y[g] = acc, or the fused Chebyshev step
// (ChebFuse; xg = x[g], the apply's input; column offset o)
template <typename T>
__device__ inline void scatter_out(const ChebFuse& cf, T* y, size_t o, int g, T acc, T xg) {
  if (!cf.b) {
    y[g] = acc;
    return;
  }
  const T r = static_cast<const T*>(cf.b)[o + g] - acc;
  T dn = T(cf.c2) * static_cast<const T*>(cf.inv)[g] * r;
  if (cf.c1 != 0.0) dn += T(cf.c1) * static_cast<T*>(cf.dv)[o + g];
  static_cast<T*>(cf.dv)[o + g] = dn;
  y[g] = xg + dn;  // y is x here
}

template <typename T>
__global__ void __launch_bounds__(BLK) k_scatter(DevSpace S, const T* __restrict__ yl, T* y, const T* xin,
                                                 int constrained, double* __restrict__ partial, long ldl, long ldy,
                                                 ChebFuse cf) {
#pragma clang fp contract(off)
  yl += (size_t)blockIdx.y * ldl;  // column of a block apply
  y += (size_t)blockIdx.y * ldy;
  if (xin) xin += (size_t)blockIdx.y * ldy;
  __shared__ double red[BLK];
  constexpr int U = 8;
  const T* __restrict__ wt = (sizeof(T) == 8) ? (const T*)S.sc_w : (const T*)S.sc_w32;
  double d = 0.0;
  // XCD x (blocks b with b % 8 == x) owns rows [x N / 8, (x + 1) N / 8) and
  // its blocks stride through them together: the rows in flight on one XCD
  // form one sliding window, so the element results they share stay in that
  // XCD's L2 (with fewer than 8 blocks the rows are split over the blocks
  // there are)
  const int nx = min((int)gridDim.x, NXCD);
  const int xcd = blockIdx.x % NXCD, jb = blockIdx.x / NXCD;
  const int nbx = gridDim.x / NXCD + (xcd < (int)(gridDim.x % NXCD) ? 1 : 0);
  const int g0 = (int)((long long)xcd * S.ndof / nx), g1 = (int)((long long)(xcd + 1) * S.ndof / nx);
  for (int g = g0 + jb * BLK + threadIdx.x; g < g1; g += nbx * BLK) {
    const int t0 = S.sc_ptr[g], t1 = S.sc_ptr[g + 1];
    T acc = T(0);
    for (int t = t0; t < t1; t += U) {
      int c[U];
      T v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) c[u] = (t + u < t1) ? S.sc_code[t + u] : 0;
#pragma unroll
      for (int u = 0; u < U; ++u) v[u] = (t + u < t1) ? yl[c[u] & 0x3fffffff] : T(0);
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (t + u < t1) {
          T w = v[u];
          if (c[u] < 0) w = w * wt[t + u];
          else if (c[u] & 0x40000000) w = -w;
          acc += w;
        }
      }
    }
    const T xg = (partial || constrained || cf.b) ? xin[g] : T(0);
    if (constrained && S.cmask[g]) acc = (T)S.diag[g] * xg;
    scatter_out(cf, y, (size_t)blockIdx.y * ldy, g, acc, xg);
    if (partial) d += (double)xg * (double)acc;
  }
  if (!partial) return;
  double tot = block_sum(d, red);
  if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// The same sums by chunks of SCH rows: a block first streams every term of
// its chunk, one per lane (coalesced code loads, independent result loads,
// the weight or sign applied exactly as above), into LDS, then each lane adds
// its row's terms from LDS in map order.  Per-term products and the order of
// the additions are those of k_scatter, so y is bit-identical; what changes
// is that the loads no longer wait on each other row by row.  Chunks are
// spread over XCDs like k_scatter's rows (XCD x owns a contiguous range).
constexpr int SCH = BLK;
static_assert(SCH == SLAB_ROWS, "one slab chunk per scatter block");
template <typename T>
__global__ void __launch_bounds__(BLK) k_scatter_chunks(DevSpace S, const T* __restrict__ yl, T* y, const T* xin,
                                                        int constrained, double* __restrict__ partial, long ldl,
                                                        long ldy, ChebFuse cf) {
#pragma clang fp contract(off)
  yl += (size_t)blockIdx.y * ldl;  // column of a block apply
  y += (size_t)blockIdx.y * ldy;
  if (xin) xin += (size_t)blockIdx.y * ldy;
  extern __shared__ unsigned char scat_raw[];
  T* vals = reinterpret_cast<T*>(scat_raw);
  __shared__ double red[BLK];
  constexpr int U = 4;
  const T* __restrict__ wt = (sizeof(T) == 8) ? (const T*)S.sc_w : (const T*)S.sc_w32;
  double d = 0.0;
  const int nch = (S.ndof + SCH - 1) / SCH;
  const int nx = min((int)gridDim.x, NXCD);
  const int xcd = blockIdx.x % NXCD, jb = blockIdx.x / NXCD;
  const int nbx = gridDim.x / NXCD + (xcd < (int)(gridDim.x % NXCD) ? 1 : 0);
  const int c0 = (int)((long long)xcd * nch / nx), c1 = (int)((long long)(xcd + 1) * nch / nx);
  const bool need_x = partial || constrained || cf.b;
  for (int c = c0 + jb; c < c1; c += nbx) {
    const int g0 = c * SCH, g1 = min(S.ndof, g0 + SCH);
    const int t0 = S.sc_ptr[g0], t1 = S.sc_ptr[g1];
    // the row's bounds, input value and elimination flag are loaded with the
    // terms, not after them
    const int g = g0 + (int)threadIdx.x;
    const int ra = g < g1 ? S.sc_ptr[g] : 0, rb = g < g1 ? S.sc_ptr[g + 1] : 0;
    const T xg = (g < g1 && need_x) ? xin[g] : T(0);
    const uint8_t cg = S.cmask[g < g1 ? g : 0];  // unconditional, as in k_scatter_slab
    for (int t = t0 + threadIdx.x; t < t1; t += U * BLK) {
      int cc[U];
      T v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) cc[u] = (t + u * BLK < t1) ? S.sc_code[t + u * BLK] : 0;
#pragma unroll
      for (int u = 0; u < U; ++u) v[u] = (t + u * BLK < t1) ? yl[cc[u] & 0x3fffffff] : T(0);
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int tt = t + u * BLK;
        if (tt < t1) {
          T w = v[u];
          if (cc[u] < 0) w = w * wt[tt];
          else if (cc[u] & 0x40000000) w = -w;
          vals[tt - t0] = w;
        }
      }
    }
    __syncthreads();
    if (g < g1) {
      const int a = ra - t0, b = rb - t0;
      T acc = T(0);
      for (int t = a; t < b; ++t) acc += vals[t];
      if (constrained && cg) acc = (T)S.diag[g] * xg;
      scatter_out(cf, y, (size_t)blockIdx.y * ldy, g, acc, xg);
      if (partial) d += (double)xg * (double)acc;
    }
    __syncthreads();  // vals is refilled by the next chunk
  }
  if (!partial) return;
  double tot = block_sum(d, red);
  if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// Scatter of entity-slab results (slab.hpp): chunk c of 256 DOFs reads its
// slab span [sl_chunk[2c], + sl_chunk[2c + 1]) into LDS with coalesced
// loads; DOF g then adds its m entries, spaced by its geometry's DOF count,
// in element order.  The slab already carries the signs and face weights,
// so there are no codes, weights or indirect loads.  A block takes SP
// consecutive chunks at a time and issues all their loads (the spans, and
// each DOF's input value and elimination flag) before waiting (the loads in
// flight, not the instructions, bound this kernel).
// Eliminated rows, partial sums and the XCD chunk ranges are those of
// k_scatter_chunks.
constexpr int SP = 4;   // chunks per step
static_assert(SP % 2 == 0 && SLAB_CHUNK_GROUP % SP == 0, "a step reads whole int4 pairs of its group");
constexpr int SU = 4;   // span entries per thread and chunk: spans <= SU * BLK
template <typename T>
__global__ void __launch_bounds__(BLK) k_scatter_slab(DevSpace S, const T* __restrict__ sl, T* y, const T* xin,
                                                      int constrained, double* __restrict__ partial, long ldl,
                                                      long ldy, ChebFuse cf) {
  sl += (size_t)blockIdx.y * ldl;  // column of a block apply
  y += (size_t)blockIdx.y * ldy;
  if (xin) xin += (size_t)blockIdx.y * ldy;
  extern __shared__ unsigned char scat_raw[];
  T* vals = reinterpret_cast<T*>(scat_raw);  // [SP][sl_chunk_max]
  __shared__ double red[BLK];
  double d = 0.0;
  const int nch = (S.ndof + SCH - 1) / SCH;
  const int ngr = (nch + SP - 1) / SP;
  const int nx = min((int)gridDim.x, NXCD);
  const int xcd = blockIdx.x % NXCD, jb = blockIdx.x / NXCD;
  const int nbx = gridDim.x / NXCD + (xcd < (int)(gridDim.x % NXCD) ? 1 : 0);
  const int q0 = (int)((long long)xcd * ngr / nx), q1 = (int)((long long)(xcd + 1) * ngr / nx);
  const int n1 = S.sl_n1, n2 = S.sl_n2, cm = S.sl_chunk_max;
  const bool need_x = partial || constrained || cf.b;
  // span bounds of the step about to run; each step loads the next step's
  // (the compiler places that load at the end of the step).  The table is
  // padded with empty chunks to whole groups, so a step's bounds are one
  // scalar load without a test
  int lo[SP], len[SP];
  auto bounds = [&](int q) {
    const int4* cb = reinterpret_cast<const int4*>(S.sl_chunk) + (size_t)q * (SP / 2);
#pragma unroll
    for (int h = 0; h < SP / 2; ++h) {
      const int4 w = cb[h];
      lo[2 * h] = w.x; len[2 * h] = w.y;
      lo[2 * h + 1] = w.z; len[2 * h + 1] = w.w;
    }
  };
  if (q0 + jb < q1) bounds(q0 + jb);
  for (int q = q0 + jb; q < q1; q += nbx) {
    int meta[SP];
    T v[SP][SU];
    T xg[SP];       // the input value at the DOF: partial sum, eliminated row
    uint8_t cg[SP];  // 1 where the DOF is eliminated
#pragma unroll
    for (int p = 0; p < SP; ++p) {
      const int c = q * SP + p;
      const int g = c * SCH + (int)threadIdx.x;
      const bool own = c < nch && g < S.ndof;
      meta[p] = own ? S.sl_meta[g] : 0;
      xg[p] = (own && need_x) ? xin[g] : T(0);
      // read unconditionally (DOF 0 stands in past the end): a flag loaded
      // under a condition is turned into a lane mask at once, and the wait
      // for it would hold back the loads after it
      cg[p] = S.cmask[own ? g : 0];
#pragma unroll
      for (int u = 0; u < SU; ++u) {
        const int t = (int)threadIdx.x + u * BLK;
        v[p][u] = (t < len[p]) ? sl[lo[p] + t] : T(0);
      }
    }
    if (q + nbx < q1) bounds(q + nbx);
#pragma unroll
    for (int p = 0; p < SP; ++p)
#pragma unroll
      for (int u = 0; u < SU; ++u) {
        const int t = (int)threadIdx.x + u * BLK;
        if (t < cm) vals[p * cm + t] = v[p][u];
      }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < SP; ++p) {
      const int g = (q * SP + p) * SCH + (int)threadIdx.x;
      if (g < S.ndof) {
        const int rel = meta[p] & 0xfffff, m = (meta[p] >> 20) & 0x3ff, dim = (int)((unsigned)meta[p] >> 30);
        const int stride = dim == 1 ? n1 : dim == 2 ? n2 : 1;
        const T* row = vals + p * cm + rel;
        T acc = T(0);
        for (int k = 0; k < m; ++k) acc += row[k * stride];
        if (constrained && cg[p]) acc = (T)S.diag[g] * xg[p];
        scatter_out(cf, y, (size_t)blockIdx.y * ldy, g, acc, xg[p]);
        if (partial) d += (double)xg[p] * (double)acc;
      }
    }
    __syncthreads();  // vals is refilled by the next step
  }
  if (!partial) return;
  double tot = block_sum(d, red);
  if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// every chunk span is loaded by one pass of SU loads per thread
bool slab_scatter_fits(const DevSpace& S) { return S.sl_chunk_max <= SU * BLK; }

static bool slab_from_env() {
  const char* v = getenv("TSEM_SLAB");
  return !(v && v[0] == '0');
}
static bool g_slab = slab_from_env();
void set_slab(bool on) { g_slab = on; }
bool slab_enabled() { return g_slab; }

// ncol > 1: a block of columns (element results ldl apart, DOF vectors ldy
// apart), with no partial sums
template <typename T>
static void slab_scatter(const DevSpace& S, const T* sl, T* y, const T* x, int constrained, double* partial,
                         hipStream_t st, int ncol = 1, long ldl = 0, long ldy = 0, const ChebFuse& cf = ChebFuse()) {
  if (ncol > 1 && partial) throw std::logic_error("block scatter takes no partial sums");
  hipLaunchKernelGGL(k_scatter_slab<T>, dim3(reduce_blocks(S.ndof), ncol), dim3(BLK),
                     (size_t)SP * S.sl_chunk_max * sizeof(T), st, S, sl, y, x, constrained, partial, ldl, ldy, cf);
}

// the chunked scatter when a chunk's terms fit in LDS, else the row kernel
template <typename T>
static void scatter_dispatch(const DevSpace& S, const T* yl, T* y, const T* x, int constrained, double* partial,
                             hipStream_t st, int ncol = 1, long ldl = 0, long ldy = 0,
                             const ChebFuse& cf = ChebFuse()) {
  if (ncol > 1 && partial) throw std::logic_error("block scatter takes no partial sums");
  const dim3 grid(reduce_blocks(S.ndof), ncol);
  size_t lds = (size_t)S.sc_chunk_max * sizeof(T);
  if (S.sc_chunk_max > 0 && lds <= 64 * 1024)
    hipLaunchKernelGGL(k_scatter_chunks<T>, grid, dim3(BLK), lds, st, S, yl, y, x, constrained, partial, ldl, ldy,
                       cf);
  else
    hipLaunchKernelGGL(k_scatter<T>, grid, dim3(BLK), 0, st, S, yl, y, x, constrained, partial, ldl, ldy, cf);
}

// ---- optional per-kernel timing with HIP events on the launching stream ----
namespace {
struct KTimer {
  bool on = false;
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  std::vector<std::pair<int, size_t>> rec;  // (kind, index of start event)
};
KTimer g_kt;
}  // namespace

void timing_enable(bool on) {
  if (on && g_kt.pool.empty()) {
    g_kt.pool.resize(2 * 40000);
    for (auto& e : g_kt.pool) (void)hipEventCreate(&e);
  }
  g_kt.on = on;
  g_kt.used = 0;
  g_kt.rec.clear();
}

void timing_read(double* out) {
  for (int k = 0; k < 2 * TK_KINDS; ++k) out[k] = 0;
  (void)hipDeviceSynchronize();
  for (auto& r : g_kt.rec) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, g_kt.pool[r.second], g_kt.pool[r.second + 1]) == hipSuccess) {
      out[2 * r.first] += ms;
      out[2 * r.first + 1] += 1;
    }
  }
}

int timing_begin(int kind, hipStream_t st) {
  if (!g_kt.on || g_kt.used + 2 > g_kt.pool.size()) return -1;
  size_t i = g_kt.used;
  g_kt.used += 2;
  (void)hipEventRecord(g_kt.pool[i], st);
  g_kt.rec.push_back({kind, i});
  return (int)i;
}
void timing_end(int h, hipStream_t st) {
  if (h >= 0) (void)hipEventRecord(g_kt.pool[h + 1], st);
}

static int apply_path_from_env() {
  const char* v = getenv("TSEM_APPLY_PATH");
  return (v && (v[0] == '1' || v[0] == '2')) ? v[0] - '0' : 0;
}
static int g_apply_path = apply_path_from_env();
void set_apply_path(int path) { g_apply_path = path; }
int apply_path() { return g_apply_path; }
bool split_path(const DevSpace& S, const OpCoef& op) {
  if (!(op.mass == 0.0 && g_apply_path == 0 && mfma_h_supported(S))) return false;
  // measured on the 32^3 mesh, element kernel + scatter per apply, split
  // kernel against the MFMA kernel with its fused gather: p = 3: 85 / 66 us,
  // p = 4: 118 / 106; from p = 5 on the split kernel is faster (p = 6: 234
  // against 319 us with an earlier, slower split kernel)
  return op.pot != 0.0 || S.K > 35 || !mfma_gx_supported(S);
}
bool block_apply_supported(const DevSpace& S, const OpCoef& op) {
  return op.ncol >= 1 && op.mass == 0.0 && op.pot == 0.0 && g_apply_path == 0 &&
         (split_path(S, op) || (S.K <= 35 && mfma_gx_supported(S)));
}

template <typename T>
static void apply_impl(const DevSpace& S, const OpCoef& op, const T* x, T* y, T* yl, double* partial, hipStream_t st) {
  const int kind = (sizeof(T) == 8) ? TK_GATHER64 : TK_GATHER32;
  if (op.ncol != 1) {
    // a block of columns: the fp16-split element kernel and the scatter,
    // each one launch over (blocks, columns)
    if (!(sizeof(T) == 4 && block_apply_supported(S, op)))
      throw std::logic_error("block apply needs the FP32 stiffness operator on the default path");
    if (!split_path(S, op)) {
      // orders up to 4: the FP32 MFMA kernel, element-major results
      launch_apply_mfma<T>(S, op, x, yl, st);
      scatter_dispatch<T>(S, yl, y, x, op.constrained, partial, st, op.ncol, op.ldl, op.ldx, op.cheb);
      return;
    }
    const bool slab = g_slab && S.sl_meta && op.Kl == 0 && slab_scatter_fits(S) && slab_h_fits(S);
    launch_apply_mfma_h(S, op, (const float*)x, (float*)yl, slab, st);
    if (slab) slab_scatter<T>(S, yl, y, x, op.constrained, partial, st, op.ncol, op.ldl, op.ldx, op.cheb);
    else scatter_dispatch<T>(S, yl, y, x, op.constrained, partial, st, op.ncol, op.ldl, op.ldx, op.cheb);
    return;
  }
  bool slab = false;
  if (sizeof(T) == 4 && split_path(S, op)) {
    // entity-slab results when the element kernel's output goes straight to
    // the scatter (full order, no potential term added to it)
    slab = g_slab && S.sl_meta && op.Kl == 0 && op.pot == 0.0 && slab_scatter_fits(S) && slab_h_fits(S);
    launch_apply_mfma_h(S, op, (const float*)x, (float*)yl, slab, st);
  } else if ((g_apply_path == 0 || g_apply_path == 2) && mfma_supported(S)) {
    launch_apply_mfma<T>(S, op, x, yl, st);
  } else {
    constexpr int NE = 8;
    size_t lds = (size_t)NE * (2 * S.K + 3 * S.Kd) * sizeof(T);
    int nb = (S.nt + NE - 1) / NE;
    int th = op.timed() ? timing_begin(kind + 1, st) : -1;
    hipLaunchKernelGGL((k_apply<T, NE>), dim3(nb), dim3(BLK), lds, st, S, op, x, yl);
    timing_end(th, st);
  }
  if (op.pot != 0.0) {
    if (!S.vpot) throw std::logic_error("potential operator requested but no potential is set");
    launch_potential<T>(S, op.pot, x, yl, op.constrained, st);
  }
  int th = op.timed() ? timing_begin(kind + 2, st) : -1;
  if (slab) slab_scatter<T>(S, yl, y, x, op.constrained, partial, st, 1, 0, 0, op.cheb);
  else scatter_dispatch<T>(S, yl, y, x, op.constrained, partial, st, 1, 0, 0, op.cheb);
  timing_end(th, st);
}

template <>
void launch_apply<double>(const DevSpace& S, const OpCoef& op, const double* x, double* y, double* yl,
                          double* partial, hipStream_t st) {
  apply_impl<double>(S, op, x, y, yl, partial, st);
}
template <>
void launch_apply<float>(const DevSpace& S, const OpCoef& op, const float* x, float* y, float* yl, double* partial,
                         hipStream_t st) {
  apply_impl<float>(S, op, x, y, yl, partial, st);
}

template <typename T>
void launch_scatter(const DevSpace& S, const T* yl, T* y, hipStream_t st) {
  scatter_dispatch<T>(S, yl, y, (const T*)nullptr, 0, (double*)nullptr, st);
}
template void launch_scatter<double>(const DevSpace&, const double*, double*, hipStream_t);
template void launch_scatter<float>(const DevSpace&, const float*, float*, hipStream_t);

// ------------------------------------------------------------- reductions ----
// reductions use a capped, n-determined grid with grid-stride loops: the
// assignment of entries to threads is fixed, so sums are reproducible
constexpr int RB_MAX = 2048;
int reduce_blocks(int n) { int b = (n + BLK - 1) / BLK; return b < RB_MAX ? b : RB_MAX; }
static inline int grid1(int n) { return (n + BLK - 1) / BLK; }

// out[v] = sum_b partial[v * nb + b], fixed tree
__global__ void __launch_bounds__(BLK) k_reduce_final(const double* __restrict__ partial, int nb, int nvals,
                                                      double* __restrict__ out) {
  __shared__ double red[BLK];
  for (int v = 0; v < nvals; ++v) {
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += BLK) s += partial[(size_t)v * nb + b];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = BLK / 2; k > 0; k >>= 1) {
      if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[v] = red[0];
    __syncthreads();
  }
}

void launch_reduce_final(const double* partial, int nb, int nvals, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(BLK), 0, st, partial, nb, nvals, out);
}

// --------------------------------------------------------------- CG steps ----
// final reductions that also advance the CG scalars
// alpha = rz / pq
__global__ void __launch_bounds__(BLK) k_final_alpha(const double* __restrict__ partial, int nb, CgScalars* sc) {
  __shared__ double red[BLK];
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += BLK) s += partial[b];
  double pq = block_sum(s, red);
  if (threadIdx.x == 0) {
    sc->pq = pq;
    sc->alpha = sc->rz / pq;
  }
}
// tmp[0..2] = r.z, |r|_1, |r|_2^2; beta = r.z_new / r.z
__global__ void __launch_bounds__(BLK) k_final_beta(const double* __restrict__ partial, int nb, int with_rz,
                                                    CgScalars* sc) {
  __shared__ double red[BLK];
  // the three sums in one pass over the partials (each in block order)
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (int b = threadIdx.x; b < nb; b += BLK) {
    s0 += partial[b];
    s1 += partial[(size_t)nb + b];
    s2 += partial[2 * (size_t)nb + b];
  }
  double v[3];
  v[0] = block_sum(s0, red);
  v[1] = block_sum(s1, red);
  v[2] = block_sum(s2, red);
  if (threadIdx.x == 0) {
    double rz = with_rz ? v[0] : sc->tmp[0];
    sc->rz_new = rz;
    sc->beta = rz / sc->rz;
    sc->rz = rz;
    sc->l1 = v[1];
    sc->l2 = v[2];
  }
}
void launch_final_alpha(const double* partial, int nb, CgScalars* sc, hipStream_t st) {
  hipLaunchKernelGGL(k_final_alpha, dim3(1), dim3(BLK), 0, st, partial, nb, sc);
}
void launch_final_beta(const double* partial, int nb, int with_rz, CgScalars* sc, hipStream_t st) {
  hipLaunchKernelGGL(k_final_beta, dim3(1), dim3(BLK), 0, st, partial, nb, with_rz, sc);
}

__global__ void k_alpha(CgScalars* sc) { sc->alpha = sc->rz / sc->pq; }
// tmp[0..2] = r.z, |r|_1, |r|_2^2 of the new residual
__global__ void k_beta(CgScalars* sc) {
  sc->rz_new = sc->tmp[0];
  sc->beta = sc->rz_new / sc->rz;
  sc->rz = sc->rz_new;
  sc->l1 = sc->tmp[1];
  sc->l2 = sc->tmp[2];
}
void launch_scalar_alpha(CgScalars* sc, hipStream_t st) { hipLaunchKernelGGL(k_alpha, dim3(1), dim3(1), 0, st, sc); }
void launch_scalar_beta(CgScalars* sc, hipStream_t st) { hipLaunchKernelGGL(k_beta, dim3(1), dim3(1), 0, st, sc); }

// x += alpha p ; r += alpha q ; z = r / diag (Jacobi) or z = r ;
// partial sums of r.z, |r|_1, |r|_2^2  (src/TSEMSolver.cpp:1846-1897).
// With pq_part, every block first sums the nb partials of p.q itself, in the
// order and with the code of k_final_alpha (so every block gets the same,
// bitwise, alpha = rz_new / pq), and block 0 records pq and alpha and moves
// rz_new into rz (rz_new is the latest r.z: k_final_beta sets both,
// k_p_update_beta only rz_new); otherwise alpha is read from sc.  pq_part
// must not overlap partial.
template <typename T>
__global__ void __launch_bounds__(BLK) k_cg_update(DevSpace S, int n, T* __restrict__ x, T* __restrict__ r,
                                                   T* __restrict__ z, const T* __restrict__ p,
                                                   const T* __restrict__ q, CgScalars* __restrict__ sc,
                                                   double* __restrict__ partial, int jac,
                                                   const double* __restrict__ pq_part, int pq_nb) {
  __shared__ double red[BLK];
  double a = 0, b = 0, c = 0;
  const T* __restrict__ dg = (sizeof(T) == 8) ? (const T*)S.invdiag : (const T*)S.invdiag32;
  // with the reduction first, the first element's loads go out before it,
  // behind its latency
  const int i0 = blockIdx.x * BLK + threadIdx.x;
  const bool pre = pq_part && i0 < n;
  const T x0 = pre ? x[i0] : T(0), p0 = pre ? p[i0] : T(0), q0 = pre ? q[i0] : T(0), r0 = pre ? r[i0] : T(0);
  const T d0 = (pre && jac) ? dg[i0] : T(0);
  double alpha_d;
  if (pq_part) {
    double s = 0.0;
    for (int k = threadIdx.x; k < pq_nb; k += BLK) s += pq_part[k];
    const double pq = block_sum(s, red);
    alpha_d = sc->rz_new / pq;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      sc->pq = pq;
      sc->alpha = alpha_d;
      sc->rz = sc->rz_new;  // the denominator of the next beta
    }
  } else {
    alpha_d = sc->alpha;
  }
  const T alpha = (T)alpha_d;
  auto step = [&](int i, T xi, T pi, T qi, T rv, T di) {
    x[i] = xi + alpha * pi;
    const T ri = rv + alpha * qi;
    r[i] = ri;
    const T zi = jac ? ri * di : ri;
    if (z) z[i] = zi;
    a += (double)ri * (double)zi;
    b += fabs((double)ri);
    c += (double)ri * (double)ri;
  };
  int i = i0;
  if (pq_part) {
    if (pre) step(i0, x0, p0, q0, r0, d0);
    i += gridDim.x * BLK;
  }
  for (; i < n; i += gridDim.x * BLK) step(i, x[i], p[i], q[i], r[i], jac ? dg[i] : T(0));
  a = block_sum(a, red);
  b = block_sum(b, red);
  c = block_sum(c, red);
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = a;
    partial[gridDim.x + blockIdx.x] = b;
    partial[2 * gridDim.x + blockIdx.x] = c;
  }
}

template <typename T>
void launch_cg_update(const DevSpace& S, int n, T* x, T* r, T* z, const T* p, const T* q, CgScalars* sc,
                      double* partial, int jac, const double* pq_part, hipStream_t st) {
  int nb = reduce_blocks(n);
  hipLaunchKernelGGL(k_cg_update<T>, dim3(nb), dim3(BLK), 0, st, S, n, x, r, z, p, q, sc, partial, jac, pq_part,
                     nb);
}
template void launch_cg_update<double>(const DevSpace&, int, double*, double*, double*, const double*, const double*,
                                       CgScalars*, double*, int, const double*, hipStream_t);
template void launch_cg_update<float>(const DevSpace&, int, float*, float*, float*, const float*, const float*,
                                      CgScalars*, double*, int, const double*, hipStream_t);

// p = beta p - z
template <typename T>
__global__ void __launch_bounds__(BLK) k_p_update(int n, T* __restrict__ p, const T* __restrict__ z,
                                                  const CgScalars* __restrict__ sc) {
  int i = blockIdx.x * BLK + threadIdx.x;
  if (i < n) p[i] = (T)sc->beta * p[i] - z[i];
}
template <typename T>
void launch_p_update(int n, T* p, const T* z, CgScalars* sc, hipStream_t st) {
  hipLaunchKernelGGL(k_p_update<T>, dim3(grid1(n)), dim3(BLK), 0, st, n, p, z, sc);
}
template void launch_p_update<double>(int, double*, const double*, CgScalars*, hipStream_t);
template void launch_p_update<float>(int, float*, const float*, CgScalars*, hipStream_t);

// k_final_beta and p = beta p - z in one kernel: every block sums the r.z
// partials [0, nb) itself, with the loop and tree of k_final_beta (so every
// block gets the same beta, bitwise, as that kernel gives), and block 0 also
// sums |r|_1 and |r|_2^2 and records rz_new, beta, l1 and l2.  sc->rz (the
// previous r.z, read by every block) is not written here: the next CG update
// with pq partials copies rz_new into it.
template <typename T>
__global__ void __launch_bounds__(BLK) k_p_update_beta(int n, T* __restrict__ p, const T* __restrict__ z,
                                                       CgScalars* __restrict__ sc,
                                                       const double* __restrict__ partial, int nb) {
  __shared__ double red[BLK];
  // the first element's loads go out before the reduction, behind its latency
  const int i0 = blockIdx.x * BLK + threadIdx.x;
  const T p0 = i0 < n ? p[i0] : T(0), z0 = i0 < n ? z[i0] : T(0);
  double s0 = 0.0;
  for (int b = threadIdx.x; b < nb; b += BLK) s0 += partial[b];
  const double rz = block_sum(s0, red);
  const double beta = rz / sc->rz;
  if (blockIdx.x == 0) {
    double s1 = 0.0, s2 = 0.0;
    for (int b = threadIdx.x; b < nb; b += BLK) {
      s1 += partial[(size_t)nb + b];
      s2 += partial[2 * (size_t)nb + b];
    }
    const double l1 = block_sum(s1, red);
    const double l2 = block_sum(s2, red);
    if (threadIdx.x == 0) {
      sc->rz_new = rz;
      sc->beta = beta;
      sc->l1 = l1;
      sc->l2 = l2;
    }
  }
  const T bt = (T)beta;
  if (i0 < n) p[i0] = bt * p0 - z0;
  for (int i = i0 + gridDim.x * BLK; i < n; i += gridDim.x * BLK) p[i] = bt * p[i] - z[i];
}
template <typename T>
void launch_p_update_beta(int n, T* p, const T* z, CgScalars* sc, const double* partial, hipStream_t st) {
  const int nb = reduce_blocks(n);
  hipLaunchKernelGGL(k_p_update_beta<T>, dim3(nb), dim3(BLK), 0, st, n, p, z, sc, partial, nb);
}
template void launch_p_update_beta<double>(int, double*, const double*, CgScalars*, const double*, hipStream_t);
template void launch_p_update_beta<float>(int, float*, const float*, CgScalars*, const double*, hipStream_t);

template <typename T>
__global__ void k_axpby(int n, T a, const T* __restrict__ x, T b, T* __restrict__ y) {
  int i = blockIdx.x * BLK + threadIdx.x;
  if (i < n) y[i] = a * x[i] + b * y[i];
}
template <typename T>
void launch_axpby(int n, T a, const T* x, T b, T* y, hipStream_t st) {
  hipLaunchKernelGGL(k_axpby<T>, dim3(grid1(n)), dim3(BLK), 0, st, n, a, x, b, y);
}
template void launch_axpby<double>(int, double, const double*, double, double*, hipStream_t);
template void launch_axpby<float>(int, float, const float*, float, float*, hipStream_t);

// zero fill as a kernel (not a memset node, so it can sit in captured graphs)
template <typename T>
__global__ void k_zero(int n, T* __restrict__ y) {
  int i = blockIdx.x * BLK + threadIdx.x;
  if (i < n) y[i] = T(0);
}
template <typename T>
void launch_zero(int n, T* y, hipStream_t st) {
  hipLaunchKernelGGL(k_zero<T>, dim3(grid1(n)), dim3(BLK), 0, st, n, y);
}
template void launch_zero<double>(int, double*, hipStream_t);
template void launch_zero<float>(int, float*, hipStream_t);

bool timing_on() { return g_kt.on; }

template <typename T>
__global__ void __launch_bounds__(BLK) k_dot(int n, const T* __restrict__ a, const T* __restrict__ b,
                                             double* __restrict__ partial) {
  __shared__ double red[BLK];
  double s = 0.0;
  for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK) s += (double)a[i] * (double)b[i];
  s = block_sum(s, red);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}
template <typename T>
void launch_dot(int n, const T* a, const T* b, double* partial, hipStream_t st) {
  hipLaunchKernelGGL(k_dot<T>, dim3(reduce_blocks(n)), dim3(BLK), 0, st, n, a, b, partial);
}
template void launch_dot<double>(int, const double*, const double*, double*, hipStream_t);
template void launch_dot<float>(int, const float*, const float*, double*, hipStream_t);

__global__ void k_cast(int n, const double* __restrict__ x, double s, float* __restrict__ y) {
  int i = blockIdx.x * BLK + threadIdx.x;
  if (i < n) y[i] = (float)(x[i] * s);
}
void launch_cast_f64_f32(int n, const double* x, double s, float* y, hipStream_t st) {
  hipLaunchKernelGGL(k_cast, dim3(grid1(n)), dim3(BLK), 0, st, n, x, s, y);
}
__global__ void k_addf(int n, const float* __restrict__ x, double s, double* __restrict__ y) {
  int i = blockIdx.x * BLK + threadIdx.x;
  if (i < n) y[i] += s * (double)x[i];
}
void launch_add_f32_f64(int n, const float* x, double s, double* y, hipStream_t st) {
  hipLaunchKernelGGL(k_addf, dim3(grid1(n)), dim3(BLK), 0, st, n, x, s, y);
}
__global__ void k_widen(int n, const float* __restrict__ x, double s, double* __restrict__ y) {
  int i = blockIdx.x * BLK + threadIdx.x;
  if (i < n) y[i] = s * (double)x[i];
}
void launch_cast_f32_f64(int n, const float* x, double s, double* y, hipStream_t st) {
  hipLaunchKernelGGL(k_widen, dim3(grid1(n)), dim3(BLK), 0, st, n, x, s, y);
}

// partial sums of |r32 + s r64|_2^2: how far an FP32 recursive residual has
// drifted from the true one it is about to be replaced by
__global__ void __launch_bounds__(BLK) k_gap(int n, const float* __restrict__ r32, const double* __restrict__ r64,
                                             double s, double* __restrict__ partial) {
  __shared__ double red[BLK];
  double a = 0;
  for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK) {
    double v = (double)r32[i] + s * r64[i];
    a += v * v;
  }
  a = block_sum(a, red);
  if (threadIdx.x == 0) partial[blockIdx.x] = a;
}
void launch_gap(int n, const float* r32, const double* r64, double s, double* partial, hipStream_t st) {
  hipLaunchKernelGGL(k_gap, dim3(reduce_blocks(n)), dim3(BLK), 0, st, n, r32, r64, s, partial);
}

// r = b - Ax ; partial |r|_2^2 and |b|_2^2
__global__ void __launch_bounds__(BLK) k_residual(int n, const double* __restrict__ b, const double* __restrict__ Ax,
                                                  double* __restrict__ r, double* __restrict__ partial) {
  __shared__ double red[BLK];
  double a = 0, c = 0;
  for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK) {
    double v = b[i] - Ax[i];
    r[i] = v;
    a += v * v;
    c += b[i] * b[i];
  }
  a = block_sum(a, red);
  c = block_sum(c, red);
  if (threadIdx.x == 0) { partial[blockIdx.x] = a; partial[gridDim.x + blockIdx.x] = c; }
}
void launch_residual(const DevSpace& S, int n, const double* b, const double* Ax, double* r, double* partial,
                     hipStream_t st) {
  (void)S;
  hipLaunchKernelGGL(k_residual, dim3(reduce_blocks(n)), dim3(BLK), 0, st, n, b, Ax, r, partial);
}

// -------------------------------------------------------------------- RHS ----
// (the load-vector kernels are in source.hip)

// physical coordinates of reference points on elements [e0, e1)
__global__ void k_qp_coords(const double* __restrict__ xe, int e0, int e1, int nq, const double* __restrict__ pts,
                            double* __restrict__ out) {
  long s = (long)blockIdx.x * BLK + threadIdx.x;
  long n = (long)(e1 - e0) * nq;
  if (s >= n) return;
  int le = (int)(s / nq), q = (int)(s % nq);
  const double* P = xe + (size_t)(e0 + le) * 12;
  double x = pts[3 * q], y = pts[3 * q + 1], z = pts[3 * q + 2], l0 = 1 - x - y - z;
  for (int d = 0; d < 3; ++d) out[s * 3 + d] = l0 * P[d] + x * P[3 + d] + y * P[6 + d] + z * P[9 + d];
}
void launch_qp_coords(const DevSpace& S, const double* xe, int e0, int e1, int nq, const double* pts, double* out,
                      hipStream_t st) {
  (void)S;
  long n = (long)(e1 - e0) * nq;
  hipLaunchKernelGGL(k_qp_coords, dim3((unsigned)((n + BLK - 1) / BLK)), dim3(BLK), 0, st, xe, e0, e1, nq, pts, out);
}

}  // namespace tsem
