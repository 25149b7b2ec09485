// Block-vector kernels for the eigen solvers: Gram matrices U^T V of tall
// blocks (fixed row partition and fixed-order sums, so results are
// reproducible), block combinations Y = S C, block residuals.  Blocks are
// column-major: column j of an n-row block starts at j * ld.
#include <hip/hip_runtime.h>

#include <stdexcept>

#include "kernels.hpp"

namespace tsem {

namespace {
constexpr int GP = 96;  // max columns of a Gram operand
constexpr int GR = 32;  // rows per LDS chunk
constexpr int CM = 64;  // max output columns of one combination launch
}  // namespace

int gram_blocks(int n) {
  int nb = (n + 4 * GR - 1) / (4 * GR);
  return nb < 512 ? (nb < 1 ? 1 : nb) : 512;
}

// partial[b p q + i q + j] = sum over the rows of block b of U[r][i] V[r][j].
// Thread (ti, tj) of the 16 x 16 grid owns outputs GT ti .. GT ti + GT - 1 by
// GT tj .. : GT = 2, 4, 6 serve operands of up to 32, 64, 96 columns, so a
// narrow Gram does no work on padding.  The next chunk of rows is loaded
// into registers while the current one is multiplied out of LDS.
template <int GT>
__global__ void __launch_bounds__(256) k_gram(int n, int p, int q, const double* __restrict__ U, long ldu,
                                              const double* __restrict__ V, long ldv, int rows_per_block,
                                              double* __restrict__ partial) {
  constexpr int W = 16 * GT;
  constexpr int LPT = GR * W / 256;  // loads per thread, operand and chunk
  __shared__ double su[GR][W + 1];
  __shared__ double sv[GR][W + 1];
  const int tid = threadIdx.x, ti = tid % 16, tj = tid / 16;
  const int r_begin = blockIdx.x * rows_per_block;
  const int r_end = min(n, r_begin + rows_per_block);
  double acc[GT][GT];
#pragma unroll
  for (int a = 0; a < GT; ++a)
#pragma unroll
    for (int c = 0; c < GT; ++c) acc[a][c] = 0.0;
  double nu[LPT], nv[LPT];
  auto fetch = [&](int r0) {
#pragma unroll
    for (int u = 0; u < LPT; ++u) {
      const int s = tid + 256 * u, r = s % GR, i = s / GR, row = r0 + r;
      const bool ok = row < r_end;
      nu[u] = (ok && i < p) ? U[(size_t)i * ldu + row] : 0.0;
      nv[u] = (ok && i < q) ? V[(size_t)i * ldv + row] : 0.0;
    }
  };
  if (r_begin < r_end) fetch(r_begin);
  for (int r0 = r_begin; r0 < r_end; r0 += GR) {
#pragma unroll
    for (int u = 0; u < LPT; ++u) {
      const int s = tid + 256 * u;
      su[s % GR][s / GR] = nu[u];
      sv[s % GR][s / GR] = nv[u];
    }
    __syncthreads();
    if (r0 + GR < r_end) fetch(r0 + GR);
    for (int r = 0; r < GR; ++r) {
      double a[GT], b[GT];
#pragma unroll
      for (int k = 0; k < GT; ++k) { a[k] = su[r][GT * ti + k]; b[k] = sv[r][GT * tj + k]; }
#pragma unroll
      for (int x = 0; x < GT; ++x)
#pragma unroll
        for (int y = 0; y < GT; ++y) acc[x][y] = fma(a[x], b[y], acc[x][y]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int x = 0; x < GT; ++x)
#pragma unroll
    for (int y = 0; y < GT; ++y) {
      int i = GT * ti + x, j = GT * tj + y;
      if (i < p && j < q) partial[(size_t)blockIdx.x * p * q + (size_t)i * q + j] = acc[x][y];
    }
}

// out[v] = sum_b partial[b nvals + v] in block order (neighbouring threads
// read neighbouring values)
__global__ void __launch_bounds__(256) k_gram_reduce(int nvals, int nb, const double* __restrict__ partial,
                                                     double* __restrict__ out) {
  int v = blockIdx.x * 256 + threadIdx.x;
  if (v >= nvals) return;
  double s = 0.0;
  for (int b = 0; b < nb; ++b) s += partial[(size_t)b * nvals + v];
  out[v] = s;
}

void launch_gram(int n, int p, int q, const double* U, long ldu, const double* V, long ldv, double* partial,
                 double* out, hipStream_t st) {
  if (p < 1 || q < 1) return;
  if (p > GP || q > GP) throw std::invalid_argument("Gram operands are limited to 96 columns");
  int nb = gram_blocks(n);
  int rpb = (n + nb - 1) / nb;
  rpb = (rpb + GR - 1) / GR * GR;
  const int w = p > q ? p : q;
  if (w <= 32)
    hipLaunchKernelGGL(k_gram<2>, dim3(nb), dim3(256), 0, st, n, p, q, U, ldu, V, ldv, rpb, partial);
  else if (w <= 64)
    hipLaunchKernelGGL(k_gram<4>, dim3(nb), dim3(256), 0, st, n, p, q, U, ldu, V, ldv, rpb, partial);
  else
    hipLaunchKernelGGL(k_gram<6>, dim3(nb), dim3(256), 0, st, n, p, q, U, ldu, V, ldv, rpb, partial);
  int nv = p * q;
  hipLaunchKernelGGL(k_gram_reduce, dim3((nv + 255) / 256), dim3(256), 0, st, nv, nb, partial, out);
}

// Y[r][j] = alpha sum_c S[r][c] C[c][j] + beta Y[r][j], C row-major p x m
// (m <= 16 MT); Y is not read when beta == 0.  On the FP64 matrix cores as
// Y^T = C^T S^T: per 16-row tile of S one wave holds the tile's p entries
// (lane (lr, lk) the entries S[r0 + lr][4 k + lk], all loads in flight at
// once), multiplies them by C^T from LDS with v_mfma_f64_16x16x4f64 and
// writes rows r0 .. r0 + 15 of each output column (value q of lane (lr, lk)
// is Y[r0 + lr][16 t + lk + 4 q]).  A wave reads its rows whole before it
// writes them and waves own disjoint rows, so Y may be columns of S itself
// (an in-place update).
constexpr int CRT = 8;  // row tiles per wave
template <int MT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) k_combine(int n, int p, int m, const double* S, long lds,
                                                 const double* __restrict__ C, double* Y, long ldy, double alpha,
                                                 double beta) {
  typedef double v4 __attribute__((ext_vector_type(4)));
  constexpr int MW = 16 * MT;
  constexpr int KS = GP / 4;  // k steps of 4 columns of S
  __shared__ double sc[GP * MW];  // C zero padded to 96 x MW
  const int pp = (p + 3) & ~3;
  for (int s = threadIdx.x; s < pp * MW; s += 256) {
    const int c = s / MW, j = s % MW;
    sc[s] = (c < p && j < m) ? C[(size_t)c * m + j] : 0.0;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, lr = lane & 15, lk = lane >> 4;
  for (int rt = 0; rt < CRT; ++rt) {
    const long r0 = ((long)blockIdx.x * 4 * CRT + (long)wave * CRT + rt) * 16;
    if (r0 >= n) break;  // wave-uniform
    const long r = r0 + lr;
    const bool ok = r < n;
    double sv[KS];
#pragma unroll
    for (int k = 0; k < KS; ++k)
      if (4 * k < pp) sv[k] = (ok && 4 * k + lk < p) ? S[(size_t)(4 * k + lk) * lds + r] : 0.0;
    v4 acc[MT];
#pragma unroll
    for (int t = 0; t < MT; ++t) acc[t] = v4{0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < KS; ++k)
      if (4 * k < pp) {
#pragma unroll
        for (int t = 0; t < MT; ++t)
          acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(sc[(4 * k + lk) * MW + 16 * t + lr], sv[k], acc[t], 0, 0, 0);
      }
    if (ok) {
#pragma unroll
      for (int t = 0; t < MT; ++t)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int j = 16 * t + lk + 4 * q;
          if (j < m) {
            double* y = Y + (size_t)j * ldy + r;
            *y = (beta != 0.0) ? alpha * acc[t][q] + beta * *y : alpha * acc[t][q];
          }
        }
    }
  }
}

void launch_combine(int n, int p, int m, const double* S, long lds, const double* C, double* Y, long ldy,
                    double alpha, double beta, hipStream_t st) {
  if (p < 1 || m < 1) return;
  if (p > GP || m > CM) throw std::invalid_argument("block combination limited to 96 x 64 coefficients");
  const long rows_per_block = 4L * CRT * 16;
  const dim3 grid((unsigned)((n + rows_per_block - 1) / rows_per_block));
  if (m <= 16)
    hipLaunchKernelGGL(k_combine<1>, grid, dim3(256), 0, st, n, p, m, S, lds, C, Y, ldy, alpha, beta);
  else if (m <= 32)
    hipLaunchKernelGGL(k_combine<2>, grid, dim3(256), 0, st, n, p, m, S, lds, C, Y, ldy, alpha, beta);
  else
    hipLaunchKernelGGL(k_combine<4>, grid, dim3(256), 0, st, n, p, m, S, lds, C, Y, ldy, alpha, beta);
}

// Y[:, j] = (float)(X[:, j] s[j]) and W[:, j] = s[j] Z[:, j] for blocks of
// ncol columns n apart (s on the device: one launch per block whatever the
// scales)
__global__ void k_cast_cols(int n, const double* __restrict__ X, const double* __restrict__ s,
                            float* __restrict__ Y) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const size_t o = (size_t)blockIdx.y * n;
  if (i < n) Y[o + i] = (float)(X[o + i] * s[blockIdx.y]);
}
__global__ void k_widen_cols(int n, const float* __restrict__ Z, const double* __restrict__ s,
                             double* __restrict__ W) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const size_t o = (size_t)blockIdx.y * n;
  if (i < n) W[o + i] = s[blockIdx.y] * (double)Z[o + i];
}
void launch_cast_cols(int n, int ncol, const double* X, const double* s, float* Y, hipStream_t st) {
  hipLaunchKernelGGL(k_cast_cols, dim3((n + 255) / 256, ncol), dim3(256), 0, st, n, X, s, Y);
}
void launch_widen_cols(int n, int ncol, const float* Z, const double* s, double* W, hipStream_t st) {
  hipLaunchKernelGGL(k_widen_cols, dim3((n + 255) / 256, ncol), dim3(256), 0, st, n, Z, s, W);
}

// R[:, j] = AX[:, j] - lam[j] BX[:, j], zero on eliminated rows
__global__ void __launch_bounds__(256) k_block_resid(int n, int m, const double* __restrict__ AX,
                                                     const double* __restrict__ BX, const double* __restrict__ lam,
                                                     const uint8_t* __restrict__ cmask, double* __restrict__ R) {
  size_t s = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (s >= (size_t)n * m) return;
  int j = (int)(s / n), r = (int)(s % n);
  R[s] = cmask[r] ? 0.0 : AX[s] - lam[j] * BX[s];
}

void launch_block_resid(int n, int m, const double* AX, const double* BX, const double* lam, const uint8_t* cmask,
                        double* R, hipStream_t st) {
  size_t tot = (size_t)n * m;
  hipLaunchKernelGGL(k_block_resid, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, n, m, AX, BX, lam, cmask,
                     R);
}

// X[:, j] = 0 on eliminated rows; optionally scaled by inv[r] (Jacobi)
__global__ void __launch_bounds__(256) k_block_mask(int n, int m, const uint8_t* __restrict__ cmask,
                                                    const double* __restrict__ inv, double* __restrict__ X) {
  size_t s = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (s >= (size_t)n * m) return;
  int r = (int)(s % n);
  X[s] = cmask[r] ? 0.0 : (inv ? X[s] * inv[r] : X[s]);
}

void launch_block_mask(int n, int m, const uint8_t* cmask, const double* inv, double* X, hipStream_t st) {
  size_t tot = (size_t)n * m;
  hipLaunchKernelGGL(k_block_mask, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, n, m, cmask, inv, X);
}

}  // namespace tsem
