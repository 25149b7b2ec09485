// Load vector on the matrix cores, and a device evaluator for the
// manufactured Poisson family the examples and benchmarks solve.
//
// k_rhs_gemm: the element load vectors of calc_rhs (src/Discretize.cpp:1396-1418),
//   b_e[i] = V_e sum_q w_q phi_i(x_q) f_e(x_q), as one FP64 GEMM per block of
//   64 elements: B (64 x K) = F (64 x nq) * W (nq x K), W[q][i] = w_q phi_i(x_q),
//   both operands staged through LDS 32 points at a time, v_mfma_f64_16x16x4f64
//   with wave w owning elements 16 w .. 16 w + 15 and every 16-mode tile.
//
// k_manufactured: u = sin(pi x) sin(pi y) sin(pi z) (1 + a g),
//   g = 1/n sum_m cos(2 pi k_m . x + phi_m), and f = -Lap u or du/dn, in one
//   pass per point.  The cosines come from one sincos per coordinate:
//   e^{i pi x} gives sin(pi x) and, squared, e^{2 pi i x}, whose integer powers
//   (k in 1..3) multiply to e^{i (2 pi k . x)}.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "handle.hpp"

namespace tsem {

namespace {
constexpr int BLK = 256;
constexpr int RE = 64;  // elements per block
constexpr int RQ = 32;  // quadrature points per LDS stage

typedef double d4 __attribute__((ext_vector_type(4)));

// blockIdx.y splits the points into gridDim.y ranges of qps (a multiple of
// RQ): with one range the block writes b_e scaled by V_e into bl, with
// several it writes its unscaled partial sums into part[y][e - e0][i] for
// k_rhs_reduce
template <int MT>
__global__ void __launch_bounds__(BLK) k_rhs_gemm(DevSpace S, int e0, int e1, int nq, int qps,
                                                  const double* __restrict__ wphi, const double* __restrict__ fq,
                                                  double* __restrict__ bl, double* __restrict__ part) {
  constexpr int NW = 16 * MT;             // modes, padded to whole tiles
  __shared__ double sF[RE][RQ + 1];       // F rows of the block's elements
  __shared__ double sW[RQ][NW + 2];       // W rows of the stage's points
  const int K = S.K;
  const int eb = e0 + blockIdx.x * RE;
  const int ne = min(RE, e1 - eb);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  d4 acc[MT];
#pragma unroll
  for (int t = 0; t < MT; ++t) acc[t] = d4{0, 0, 0, 0};
  const double* frow = fq + (size_t)(eb - e0) * nq;
  const int qa = blockIdx.y * qps, qz = min(nq, qa + qps);
  for (int q0 = qa; q0 < qz; q0 += RQ) {
    for (int s = tid; s < RE * RQ; s += BLK) {
      const int le = s / RQ, qq = s - le * RQ, q = q0 + qq;
      sF[le][qq] = (le < ne && q < qz) ? frow[(size_t)le * nq + q] : 0.0;
    }
    for (int s = tid; s < RQ * NW; s += BLK) {
      const int qq = s / NW, i = s - qq * NW, q = q0 + qq;
      sW[qq][i] = (q < qz && i < K) ? wphi[(size_t)q * K + i] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < RQ / 4; ++ks) {
      const double a = sF[16 * wave + lr][4 * ks + lk];
#pragma unroll
      for (int t = 0; t < MT; ++t) {
        const double b = sW[4 * ks + lk][16 * t + lr];
        acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[t], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  // accumulator value j of lane l is C[l / 16 + 4 j][l % 16]: element
  // 16 wave + lk + 4 j, mode 16 t + lr
#pragma unroll
  for (int t = 0; t < MT; ++t)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int le = 16 * wave + lk + 4 * j, i = 16 * t + lr;
      if (le < ne && i < K) {
        const int e = eb + le;
        if (gridDim.y == 1) bl[(size_t)e * K + i] = acc[t][j] * (S.sixv[e] / 6.0);
        else part[((size_t)blockIdx.y * (e1 - e0) + (e - e0)) * K + i] = acc[t][j];
      }
    }
}

// b_e[i] = V_e sum over the point ranges, in range order
__global__ void k_rhs_reduce(DevSpace S, int e0, int e1, int ns, const double* __restrict__ part,
                             double* __restrict__ bl) {
  const long s = (long)blockIdx.x * BLK + threadIdx.x;
  const long n = (long)(e1 - e0) * S.K;
  if (s >= n) return;
  double acc = 0;
  for (int y = 0; y < ns; ++y) acc += part[(size_t)y * n + s];
  const int e = e0 + (int)(s / S.K);
  bl[(size_t)e0 * S.K + s] = acc * (S.sixv[e] / 6.0);
}

// the same sums one (element, mode) per thread, for orders past the GEMM's tiles
__global__ void __launch_bounds__(BLK) k_rhs_local(DevSpace S, int e0, int e1, int nq, const double* __restrict__ wphi,
                                                   const double* __restrict__ fq, double* __restrict__ bl) {
  long s = (long)blockIdx.x * BLK + threadIdx.x;
  long n = (long)(e1 - e0) * S.K;
  if (s >= n) return;
  int le = (int)(s / S.K), i = (int)(s % S.K);
  int e = e0 + le;
  const double* f = fq + (size_t)le * nq;
  double acc = 0;
  for (int q = 0; q < nq; ++q) acc += wphi[(size_t)q * S.K + i] * f[q];
  bl[(size_t)e * S.K + i] = acc * (S.sixv[e] / 6.0);
}

struct Cx {
  double re, im;
};
__device__ inline Cx mul(Cx a, Cx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
// e[k] for a wave number k in 0..3 by selects: an array indexed at run time
// would live in scratch memory
__device__ inline Cx pick(const Cx (&e)[4], int k) {
  const Cx a = (k & 1) ? e[1] : e[0], b = (k & 1) ? e[3] : e[2];
  return (k & 2) ? b : a;
}

struct Phases {
  double c[16], s[16];  // cos and sin of the phases
};

__global__ void __launch_bounds__(BLK) k_manufactured(tsem_manufactured_def P, Phases ph, int what, long n,
                                                      const double* __restrict__ pts, const double* __restrict__ nrm,
                                                      double* __restrict__ out) {
  const long p = (long)blockIdx.x * BLK + threadIdx.x;
  if (p >= n) return;
  const double pi = 3.14159265358979323846;
  double s[3], c[3];
  Cx e[3][4];  // e[d][k] = e^{2 pi i k x_d}, k = 0..3
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    sincospi(pts[3 * p + d], &s[d], &c[d]);
    const Cx e1 = {c[d] * c[d] - s[d] * s[d], 2 * s[d] * c[d]};
    e[d][0] = {1.0, 0.0};
    e[d][1] = e1;
    e[d][2] = mul(e1, e1);
    e[d][3] = mul(e[d][2], e1);
  }
  const double S = s[0] * s[1] * s[2];
  const double gS[3] = {pi * c[0] * s[1] * s[2], pi * s[0] * c[1] * s[2], pi * s[0] * s[1] * c[2]};
  double G = 0, gG[3] = {0, 0, 0}, LG = 0;
  const double inv_n = 1.0 / P.nterms;
  for (int m = 0; m < P.nterms; ++m) {
    const int k0 = (int)P.k[m][0], k1 = (int)P.k[m][1], k2 = (int)P.k[m][2];
    const Cx z = mul(mul(Cx{ph.c[m], ph.s[m]}, pick(e[0], k0)), mul(pick(e[1], k1), pick(e[2], k2)));  // e^{i (2 pi k.x + phi)}
    G += z.re * inv_n;
    gG[0] -= 2 * pi * P.k[m][0] * z.im * inv_n;
    gG[1] -= 2 * pi * P.k[m][1] * z.im * inv_n;
    gG[2] -= 2 * pi * P.k[m][2] * z.im * inv_n;
    LG -= 4 * pi * pi * (P.k[m][0] * P.k[m][0] + P.k[m][1] * P.k[m][1] + P.k[m][2] * P.k[m][2]) * z.re * inv_n;
  }
  const double a = P.amp;
  double v;
  if (what == 0) {
    v = S * (1 + a * G);
  } else if (what == 1) {
    const double lap = -3 * pi * pi * S * (1 + a * G) + 2 * a * (gS[0] * gG[0] + gS[1] * gG[1] + gS[2] * gG[2]) +
                       a * S * LG;
    v = -lap;
  } else {
    v = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) v += (gS[d] * (1 + a * G) + a * S * gG[d]) * nrm[3 * p + d];
  }
  out[p] = v;
}
}  // namespace

int rhs_splits(const DevSpace& S, int ne, int nq) {
  if (S.K > 176) return 1;
  const int nb = (ne + RE - 1) / RE;
  // about two blocks per CU of the 256, each point range at least 8 stages
  int ns = (512 + nb - 1) / nb;
  ns = std::min(ns, std::max(1, nq / (8 * RQ)));
  return std::max(1, std::min(ns, 32));
}

void launch_rhs_local(const DevSpace& S, int e0, int e1, int nq, const double* wphi, const double* fq, double* bl,
                      double* part, hipStream_t st) {
  const int mt = (S.K + 15) / 16;
  const int ns = part ? rhs_splits(S, e1 - e0, nq) : 1;
  const int qps = ((nq + ns - 1) / ns + RQ - 1) / RQ * RQ;
  const int nsu = (nq + qps - 1) / qps;  // ranges actually used
  const dim3 grid((unsigned)((e1 - e0 + RE - 1) / RE), (unsigned)nsu);
  bool done = true;
  switch (mt) {
#define RHS_CASE(MTV)                                                                                    \
  hipLaunchKernelGGL(k_rhs_gemm<MTV>, grid, dim3(BLK), 0, st, S, e0, e1, nq, qps, wphi, fq, bl, part); \
  break;
    case 1: RHS_CASE(1)
    case 2: RHS_CASE(2)
    case 3: RHS_CASE(3)
    case 4: RHS_CASE(4)
    case 5: case 6: RHS_CASE(6)
    case 7: case 8: RHS_CASE(8)
    case 9: case 10: case 11: RHS_CASE(11)
#undef RHS_CASE
    default: done = false;
  }
  if (done) {
    if (nsu > 1) {
      const long n = (long)(e1 - e0) * S.K;
      hipLaunchKernelGGL(k_rhs_reduce, dim3((unsigned)((n + BLK - 1) / BLK)), dim3(BLK), 0, st, S, e0, e1, nsu, part,
                         bl);
    }
    return;
  }
  const long n = (long)(e1 - e0) * S.K;
  hipLaunchKernelGGL(k_rhs_local, dim3((unsigned)((n + BLK - 1) / BLK)), dim3(BLK), 0, st, S, e0, e1, nq, wphi, fq,
                     bl);
}

}  // namespace tsem

extern "C" int tsem_manufactured(const tsem_manufactured_def* def, int what, int64_t n, const double* pts,
                                 const double* normals, double* out, void* stream) {
  API_BEGIN
  if (!def) throw std::invalid_argument("null argument");
  if (n == 0) return TSEM_OK;  // no points: nothing to read or write
  if (!pts || !out) throw std::invalid_argument("null argument");
  if (def->nterms < 1 || def->nterms > 16) throw std::invalid_argument("nterms must be in 1..16");
  for (int m = 0; m < def->nterms; ++m)
    for (int d = 0; d < 3; ++d)
      if (!(def->k[m][d] >= 0 && def->k[m][d] <= 3 && def->k[m][d] == (double)(int)def->k[m][d]))
        throw std::invalid_argument("wave numbers must be integers in 0..3");
  if (what < 0 || what > 2) throw std::invalid_argument("what must be 0 (u), 1 (f = -Lap u) or 2 (du/dn)");
  if (what == 2 && !normals) throw std::invalid_argument("du/dn needs the normals");
  if (n < 0) throw std::invalid_argument("negative point count");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const long nb = (long)((n + 255) / 256);
  tsem::Phases ph;
  for (int m = 0; m < def->nterms; ++m) {
    ph.c[m] = std::cos(def->phi[m]);
    ph.s[m] = std::sin(def->phi[m]);
  }
  hipLaunchKernelGGL(tsem::k_manufactured, dim3((unsigned)nb), dim3(256), 0, st, *def, ph, what, (long)n, pts,
                     normals, out);
  HIPCHK(hipGetLastError());
  API_END
}
