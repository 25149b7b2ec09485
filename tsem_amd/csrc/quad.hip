// Work at the 3-D quadrature points of a space: values of a DOF vector
// (TSEM::calc_val_qp_onElement, src/PointInfo.cpp:60-91), the density of
// occupied states (TSEM::calc_density, src/PointInfo.cpp:2012-2022),
// weighted integrals and low-order moments, the L2 error
// (TSEM::calc_l2_error, src/CalcError.cpp:20-126), and LDA exchange and
// correlation.  Point values of an element come from its gathered
// coefficients: u(x_q) = sum_i phi_i(x_q) xl[e][i].  Integrals over an
// element are vol_e sum_q w_q (.), the weights summing to 1.
#include <hip/hip_runtime.h>

#include <cmath>

#include "kernels.hpp"

namespace tsem {

namespace {
constexpr int QB = 256;
constexpr int IB = 512;  // max blocks of the integral kernels
constexpr int MOM = 10;  // moments {1, x, y, z, xx, xy, xz, yy, yz, zz}

__device__ inline double block_reduce(double v, double* red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int k = QB / 2; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  double r = red[0];
  __syncthreads();
  return r;
}

__device__ inline void point(const DevSpace& S, int e, int q, double* x) {
  const double* P = S.xe + (size_t)e * 12;
  double a = S.pts[3 * q], b = S.pts[3 * q + 1], c = S.pts[3 * q + 2], l0 = 1 - a - b - c;
  for (int d = 0; d < 3; ++d) x[d] = l0 * P[d] + a * P[3 + d] + b * P[6 + d] + c * P[9 + d];
}
}  // namespace

// one block per element, threads over the points
__global__ void __launch_bounds__(QB) k_qp_values(DevSpace S, int e0, const double* __restrict__ xl,
                                                  double* __restrict__ out) {
  extern __shared__ double xs[];
  const int e = e0 + blockIdx.x, K = S.K, nq = S.nq;
  for (int i = threadIdx.x; i < K; i += QB) xs[i] = xl[(size_t)e * K + i];
  __syncthreads();
  for (int q = threadIdx.x; q < nq; q += QB) {
    double u = 0.0;
    for (int i = 0; i < K; ++i) u += S.phiT[(size_t)i * nq + q] * xs[i];
    out[(size_t)blockIdx.x * nq + q] = u;
  }
}

void launch_qp_values(const DevSpace& S, int e0, int e1, const double* xl, double* out, hipStream_t st) {
  if (e1 <= e0) return;
  hipLaunchKernelGGL(k_qp_values, dim3(e1 - e0), dim3(QB), (size_t)S.K * 8, st, S, e0, xl, out);
}

__global__ void __launch_bounds__(QB) k_density(DevSpace S, int nk, const double* __restrict__ xl,
                                                const double* __restrict__ occ, double* __restrict__ rho) {
  extern __shared__ double xs[];  // nk x K
  const int e = blockIdx.x, K = S.K, nq = S.nq;
  for (int s = threadIdx.x; s < nk * K; s += QB) {
    int k = s / K, i = s % K;
    xs[s] = xl[((size_t)k * S.nt + e) * K + i];
  }
  __syncthreads();
  for (int q = threadIdx.x; q < nq; q += QB) {
    double r = 0.0;
    for (int k = 0; k < nk; ++k) {
      double u = 0.0;
      for (int i = 0; i < K; ++i) u += S.phiT[(size_t)i * nq + q] * xs[k * K + i];
      r += occ[k] * u * u;
    }
    rho[(size_t)e * nq + q] = r;
  }
}

void launch_density(const DevSpace& S, int nk, const double* xl, const double* occ, double* rho, hipStream_t st) {
  hipLaunchKernelGGL(k_density, dim3(S.nt), dim3(QB), (size_t)nk * S.K * 8, st, S, nk, xl, occ, rho);
}

int integrate_blocks(const DevSpace& S) { return S.nt < IB ? S.nt : IB; }

// partial[v * nb + b]: block b sums a fixed contiguous range of elements
__global__ void __launch_bounds__(QB) k_integrate(DevSpace S, const double* __restrict__ f,
                                                  const double* __restrict__ g, int nmom,
                                                  double* __restrict__ partial) {
  __shared__ double red[QB];
  const int nb = gridDim.x, nq = S.nq;
  const int per = (S.nt + nb - 1) / nb;
  const int ea = blockIdx.x * per, eb = min(S.nt, ea + per);
  double acc[MOM];
  for (int v = 0; v < MOM; ++v) acc[v] = 0.0;
  for (int e = ea; e < eb; ++e) {
    const double vol = S.sixv[e] / 6.0;
    for (int q = threadIdx.x; q < nq; q += QB) {
      size_t s = (size_t)e * nq + q;
      double w = vol * S.wq[q] * f[s] * (g ? g[s] : 1.0);
      acc[0] += w;
      if (nmom > 1) {
        double x[3];
        point(S, e, q, x);
        acc[1] += w * x[0]; acc[2] += w * x[1]; acc[3] += w * x[2];
        acc[4] += w * x[0] * x[0]; acc[5] += w * x[0] * x[1]; acc[6] += w * x[0] * x[2];
        acc[7] += w * x[1] * x[1]; acc[8] += w * x[1] * x[2]; acc[9] += w * x[2] * x[2];
      }
    }
  }
  for (int v = 0; v < nmom; ++v) {
    double t = block_reduce(acc[v], red);
    if (threadIdx.x == 0) partial[(size_t)v * nb + blockIdx.x] = t;
  }
}

// out[v] = sum_b partial[v nb + b] in block order
__global__ void k_sum_partials(const double* __restrict__ partial, int nb, int nv, double* __restrict__ out) {
  int v = threadIdx.x;
  if (v >= nv) return;
  double s = 0.0;
  for (int b = 0; b < nb; ++b) s += partial[(size_t)v * nb + b];
  out[v] = s;
}

void launch_integrate(const DevSpace& S, const double* f, const double* g, int nmom, double* partial, double* out,
                      hipStream_t st) {
  int nb = integrate_blocks(S);
  hipLaunchKernelGGL(k_integrate, dim3(nb), dim3(QB), 0, st, S, f, g, nmom, partial);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(64), 0, st, partial, nb, nmom, out);
}

__global__ void __launch_bounds__(QB) k_l2_chunk(DevSpace S, int e0, int e1, const double* __restrict__ xl,
                                                 const double* __restrict__ u, double* __restrict__ partial) {
  extern __shared__ double sm[];
  double* red = sm;        // QB
  double* xs = sm + QB;    // K
  const int nb = gridDim.x, nq = S.nq, K = S.K;
  const int n = e1 - e0, per = (n + nb - 1) / nb;
  const int ea = e0 + blockIdx.x * per, eb = min(e1, ea + per);
  double acc = 0.0;
  for (int e = ea; e < eb; ++e) {
    __syncthreads();
    for (int i = threadIdx.x; i < K; i += QB) xs[i] = xl[(size_t)e * K + i];
    __syncthreads();
    const double vol = S.sixv[e] / 6.0;
    for (int q = threadIdx.x; q < nq; q += QB) {
      double uh = 0.0;
      for (int i = 0; i < K; ++i) uh += S.phiT[(size_t)i * nq + q] * xs[i];
      double d = uh - u[(size_t)(e - e0) * nq + q];
      acc += vol * S.wq[q] * d * d;
    }
  }
  __syncthreads();
  double t = block_reduce(acc, red);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

void launch_l2_chunk(const DevSpace& S, int e0, int e1, const double* xl, const double* u, double* partial,
                     double* out, hipStream_t st) {
  int n = e1 - e0;
  if (n <= 0) return;
  int nb = n < IB ? n : IB;
  hipLaunchKernelGGL(k_l2_chunk, dim3(nb), dim3(QB), (size_t)(QB + S.K) * 8, st, S, e0, e1, xl, u, partial);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(64), 0, st, partial, nb, 1, out);
}

// sum_e V_e sum_q w_q |grad u_h - g|^2 over elements [e0, e1): grad u_h =
// J^-T sum_i c_i grad_ref phi_i at the point, J = [P1-P0, P2-P0, P3-P0] the
// element map of the quadrature coordinates (calc_l2_error_gradient,
// src/CalcError.cpp:199-267)
__global__ void __launch_bounds__(QB) k_h1_chunk(DevSpace S, int e0, int e1, const double* __restrict__ xl,
                                                 const double* __restrict__ g, double* __restrict__ partial) {
  extern __shared__ double sm[];
  double* red = sm;        // QB
  double* xs = sm + QB;    // K
  const int nb = gridDim.x, nq = S.nq, K = S.K;
  const int n = e1 - e0, per = (n + nb - 1) / nb;
  const int ea = e0 + blockIdx.x * per, eb = min(e1, ea + per);
  double acc = 0.0;
  for (int e = ea; e < eb; ++e) {
    __syncthreads();
    for (int i = threadIdx.x; i < K; i += QB) xs[i] = xl[(size_t)e * K + i];
    __syncthreads();
    const double* P = S.xe + (size_t)e * 12;
    double J[3][3];  // J[d][k] = d x_d / d xref_k
    for (int d = 0; d < 3; ++d)
      for (int k = 0; k < 3; ++k) J[d][k] = P[3 * (k + 1) + d] - P[d];
    const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                       J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                       J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    double Ji[3][3];  // J^-1: Ji[k][d] = d xref_k / d x_d
    Ji[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
    Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
    Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
    Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
    Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
    Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
    Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
    Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
    Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
    const double vol = S.sixv[e] / 6.0;
    for (int q = threadIdx.x; q < nq; q += QB) {
      double gr[3] = {0.0, 0.0, 0.0};
      const double* dp = S.dphi + (size_t)q * K * 3;
      for (int i = 0; i < K; ++i)
        for (int k = 0; k < 3; ++k) gr[k] += dp[3 * i + k] * xs[i];
      const double* ge = g + ((size_t)(e - e0) * nq + q) * 3;
      double s = 0.0;
      for (int d = 0; d < 3; ++d) {
        const double gd = Ji[0][d] * gr[0] + Ji[1][d] * gr[1] + Ji[2][d] * gr[2] - ge[d];
        s += gd * gd;
      }
      acc += vol * S.wq[q] * s;
    }
  }
  __syncthreads();
  double t = block_reduce(acc, red);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

void launch_h1_chunk(const DevSpace& S, int e0, int e1, const double* xl, const double* g, double* partial,
                     double* out, hipStream_t st) {
  int n = e1 - e0;
  if (n <= 0) return;
  int nb = n < IB ? n : IB;
  hipLaunchKernelGGL(k_h1_chunk, dim3(nb), dim3(QB), (size_t)(QB + S.K) * 8, st, S, e0, e1, xl, g, partial);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(64), 0, st, partial, nb, 1, out);
}

// Perdew & Zunger, Phys. Rev. B 23, 5048 (1981), unpolarized fit to the
// Ceperley-Alder correlation energy; Slater exchange eps_x = -(3/4)(3 rho/pi)^(1/3).
__global__ void __launch_bounds__(QB) k_lda_pz(long n, const double* __restrict__ rho, double* __restrict__ vxc,
                                               double* __restrict__ exc) {
  long s = (long)blockIdx.x * QB + threadIdx.x;
  if (s >= n) return;
  const double r = rho[s];
  if (!(r > 1e-30)) {
    if (vxc) vxc[s] = 0.0;
    if (exc) exc[s] = 0.0;
    return;
  }
  const double PI = 3.14159265358979323846;
  const double kx = cbrt(3.0 * r / PI);
  const double ex = -0.75 * kx, vx = -kx;
  const double rs = cbrt(3.0 / (4.0 * PI * r));
  double ec, vc;
  if (rs >= 1.0) {
    const double g = -0.1423, b1 = 1.0529, b2 = 0.3334;
    const double sr = sqrt(rs), den = 1.0 + b1 * sr + b2 * rs;
    ec = g / den;
    vc = ec * (1.0 + 7.0 / 6.0 * b1 * sr + 4.0 / 3.0 * b2 * rs) / den;
  } else {
    const double A = 0.0311, B = -0.048, C = 0.0020, D = -0.0116;
    const double lr = log(rs);
    ec = A * lr + B + C * rs * lr + D * rs;
    vc = A * lr + (B - A / 3.0) + 2.0 / 3.0 * C * rs * lr + (2.0 * D - C) / 3.0 * rs;
  }
  if (vxc) vxc[s] = vx + vc;
  if (exc) exc[s] = r * (ex + ec);
}

void launch_lda_pz(long n, const double* rho, double* vxc, double* exc, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_lda_pz, dim3((unsigned)((n + QB - 1) / QB)), dim3(QB), 0, st, n, rho, vxc, exc);
}

// Assembled diagonal of the stiffness operator from the transposed gather
// map: row g's terms come grouped by element, and element e contributes
// sum_{u,v} w_u w_v K_e[i_u][i_v] over its terms, K_e = sum_ab Ghat_ab(e)
// grad2[ab] (grad2: 9 x K x K, the reference gradient Gram matrices).
// Also the inverse diagonal in FP64 and FP32.
__global__ void __launch_bounds__(QB) k_diagonal(DevSpace S, const double* __restrict__ grad2,
                                                 double* __restrict__ diag, double* __restrict__ inv,
                                                 float* __restrict__ inv32) {
  const int g = blockIdx.x * QB + threadIdx.x;
  if (g >= S.ndof) return;
  const int K = S.K;
  const size_t KK = (size_t)K * K;
  const int t0 = S.sc_ptr[g], t1 = S.sc_ptr[g + 1];
  double acc = 0.0;
  for (int a = t0; a < t1;) {
    const int e = (S.sc_code[a] & 0x3fffffff) / K;
    int b = a + 1;
    while (b < t1 && (S.sc_code[b] & 0x3fffffff) / K == e) ++b;
    const double* gh = S.ghat + 6 * (size_t)e;
    const double G[9] = {gh[0], gh[3], gh[4], gh[3], gh[1], gh[5], gh[4], gh[5], gh[2]};
    double s = 0.0;
    for (int u = a; u < b; ++u) {
      int cu = S.sc_code[u], iu = (cu & 0x3fffffff) % K;
      double wu = cu < 0 ? S.sc_w[u] : ((cu & 0x40000000) ? -1.0 : 1.0);
      for (int v = a; v < b; ++v) {
        int cv = S.sc_code[v], iv = (cv & 0x3fffffff) % K;
        double wv = cv < 0 ? S.sc_w[v] : ((cv & 0x40000000) ? -1.0 : 1.0);
        double k = 0.0;
        for (int ab = 0; ab < 9; ++ab) k += G[ab] * grad2[ab * KK + (size_t)iu * K + iv];
        s += wu * wv * k;
      }
    }
    acc += s;
    a = b;
  }
  diag[g] = acc;
  inv[g] = 1.0 / acc;
  inv32[g] = (float)(1.0 / acc);
}

void launch_diagonal(const DevSpace& S, const double* grad2, double* diag, double* inv, float* inv32, hipStream_t st) {
  hipLaunchKernelGGL(k_diagonal, dim3((S.ndof + QB - 1) / QB), dim3(QB), 0, st, S, grad2, diag, inv, inv32);
}

// out[p] = sum_i phi[p][i] xl[e_p][i] in mode order; NaN where e_p < 0
__global__ void __launch_bounds__(QB) k_point_values(int K, long n, const int32_t* __restrict__ ep,
                                                     const double* __restrict__ phi, const double* __restrict__ xl,
                                                     double* __restrict__ out) {
  const long p = (long)blockIdx.x * QB + threadIdx.x;
  if (p >= n) return;
  const int e = ep[p];
  if (e < 0) {
    out[p] = __builtin_nan("");
    return;
  }
  double s = 0.0;
  for (int i = 0; i < K; ++i) s += phi[p * K + i] * xl[(size_t)e * K + i];
  out[p] = s;
}

void launch_point_values(int K, long n, const int32_t* ep, const double* phi, const double* xl, double* out,
                         hipStream_t st) {
  hipLaunchKernelGGL(k_point_values, dim3((unsigned)((n + QB - 1) / QB)), dim3(QB), 0, st, K, n, ep, phi, xl, out);
}

}  // namespace tsem
