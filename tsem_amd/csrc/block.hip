// Block-vector kernels for the eigen solvers: Gram matrices U^T V of tall
// blocks (fixed row partition and fixed-order sums, so results are
// reproducible), block combinations Y = S C, block residuals.  Blocks are
// column-major: column j of an n-row block starts at j * ld.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace tsem {

namespace {
constexpr int GP = 96;  // max columns of a Gram operand
constexpr int GR = 32;  // rows per LDS chunk
constexpr int GT = 6;   // outputs per thread in each direction (16 x 6 = 96)
constexpr int CM = 32;  // max output columns of one combination launch
}  // namespace

int gram_blocks(int n) {
  int nb = (n + 4 * GR - 1) / (4 * GR);
  return nb < 512 ? (nb < 1 ? 1 : nb) : 512;
}

// partial[(i q + j) nb + b] = sum over the rows of block b of U[r][i] V[r][j]
__global__ void __launch_bounds__(256) k_gram(int n, int p, int q, const double* __restrict__ U, long ldu,
                                              const double* __restrict__ V, long ldv, int rows_per_block,
                                              double* __restrict__ partial) {
  __shared__ double su[GR][GP];
  __shared__ double sv[GR][GP];
  const int tid = threadIdx.x, ti = tid % 16, tj = tid / 16;
  const int nb = gridDim.x;
  const int r_begin = blockIdx.x * rows_per_block;
  const int r_end = min(n, r_begin + rows_per_block);
  double acc[GT][GT];
#pragma unroll
  for (int a = 0; a < GT; ++a)
#pragma unroll
    for (int c = 0; c < GT; ++c) acc[a][c] = 0.0;
  for (int r0 = r_begin; r0 < r_end; r0 += GR) {
    for (int s = tid; s < GR * GP; s += 256) {
      int r = s % GR, i = s / GR, row = r0 + r;
      bool ok = row < r_end;
      su[r][i] = (ok && i < p) ? U[(size_t)i * ldu + row] : 0.0;
      sv[r][i] = (ok && i < q) ? V[(size_t)i * ldv + row] : 0.0;
    }
    __syncthreads();
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
      if (i < p && j < q) partial[((size_t)i * q + j) * nb + blockIdx.x] = acc[x][y];
    }
}

// out[v] = sum_b partial[v nb + b] in block order
__global__ void __launch_bounds__(256) k_gram_reduce(int nvals, int nb, const double* __restrict__ partial,
                                                     double* __restrict__ out) {
  int v = blockIdx.x * 256 + threadIdx.x;
  if (v >= nvals) return;
  double s = 0.0;
  for (int b = 0; b < nb; ++b) s += partial[(size_t)v * nb + b];
  out[v] = s;
}

void launch_gram(int n, int p, int q, const double* U, long ldu, const double* V, long ldv, double* partial,
                 double* out, hipStream_t st) {
  if (p < 1 || q < 1) return;
  int nb = gram_blocks(n);
  int rpb = (n + nb - 1) / nb;
  rpb = (rpb + GR - 1) / GR * GR;
  hipLaunchKernelGGL(k_gram, dim3(nb), dim3(256), 0, st, n, p, q, U, ldu, V, ldv, rpb, partial);
  int nv = p * q;
  hipLaunchKernelGGL(k_gram_reduce, dim3((nv + 255) / 256), dim3(256), 0, st, nv, nb, partial, out);
}

// Y[r][j] = alpha sum_c S[r][c] C[c][j] + beta Y[r][j], C row-major p x m
// (m <= CM); Y is not read when beta == 0
__global__ void __launch_bounds__(256) k_combine(int n, int p, int m, const double* __restrict__ S, long lds,
                                                 const double* __restrict__ C, double* __restrict__ Y, long ldy,
                                                 double alpha, double beta) {
  __shared__ double sc[GP * CM];
  for (int s = threadIdx.x; s < p * m; s += 256) sc[s] = C[s];
  __syncthreads();
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= n) return;
  double acc[CM];
#pragma unroll
  for (int j = 0; j < CM; ++j) acc[j] = 0.0;
  for (int c = 0; c < p; ++c) {
    double s = S[(size_t)c * lds + r];
#pragma unroll
    for (int j = 0; j < CM; ++j)
      if (j < m) acc[j] = fma(s, sc[c * m + j], acc[j]);
  }
#pragma unroll
  for (int j = 0; j < CM; ++j)
    if (j < m) {
      double* y = Y + (size_t)j * ldy + r;
      *y = (beta != 0.0) ? alpha * acc[j] + beta * *y : alpha * acc[j];
    }
}

void launch_combine(int n, int p, int m, const double* S, long lds, const double* C, double* Y, long ldy,
                    double alpha, double beta, hipStream_t st) {
  if (p < 1 || m < 1) return;
  hipLaunchKernelGGL(k_combine, dim3((n + 255) / 256), dim3(256), 0, st, n, p, m, S, lds, C, Y, ldy, alpha, beta);
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
