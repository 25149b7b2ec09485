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
// GT tj .. : GT = 4, 6 serve operands of up to 64, 96 columns (up to 32, the
// matrix-core kernel below).  The next chunk of rows is loaded
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

// The same partial sums on the FP64 matrix cores, for operands of up to 32
// columns: D (16 TI x 16 TJ) = U^T V over the block's rows, 16 rows a step.
// Lane (lr, lk) loads rows r + 4 lk .. r + 4 lk + 3 of column 16 ti + lr of U
// and of column 16 tj + lr of V (the four lanes of a column read one 128-byte
// run), and step m of v_mfma_f64_16x16x4f64 takes row r + 4 lk + m as k index
// lk of both operands.  Wave w takes the block's steps w, w + 4, ..., with the
// next step's rows loaded while the current one is multiplied; the four
// waves' tiles are added in wave order through LDS.  The sums have a fixed
// order, so results are reproducible.  The vector kernel above spends its
// time reading operands out of LDS (a 16 x 16 thread grid reads 2 GT values a
// row for GT^2 FMAs); here each loaded value feeds 4 TJ (or 4 TI) MFMA lanes
// straight from registers.
template <int TI, int TJ>
__global__ void __launch_bounds__(256) k_gram_mfma(int n, int p, int q, const double* __restrict__ U, long ldu,
                                                   const double* __restrict__ V, long ldv, int rows_per_block,
                                                   double* __restrict__ partial) {
  typedef double v4 __attribute__((ext_vector_type(4)));
  __shared__ double red[4][16 * TI][16 * TJ + 1];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, lr = lane & 15, lk = lane >> 4;
  const int r_begin = blockIdx.x * rows_per_block;
  const int r_end = min(n, r_begin + rows_per_block);
  v4 acc[TI][TJ];
#pragma unroll
  for (int a = 0; a < TI; ++a)
#pragma unroll
    for (int c = 0; c < TJ; ++c) acc[a][c] = v4{0, 0, 0, 0};
  double a[TI][4], b[TJ][4], na[TI][4], nb[TJ][4];
  auto fetch = [&](int r, double (&fa)[TI][4], double (&fb)[TJ][4]) {
#pragma unroll
    for (int ti = 0; ti < TI; ++ti) {
      const int c = 16 * ti + lr;
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        const int row = r + 4 * lk + m;
        fa[ti][m] = (c < p && row < r_end) ? U[(size_t)c * ldu + row] : 0.0;
      }
    }
#pragma unroll
    for (int tj = 0; tj < TJ; ++tj) {
      const int c = 16 * tj + lr;
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        const int row = r + 4 * lk + m;
        fb[tj][m] = (c < q && row < r_end) ? V[(size_t)c * ldv + row] : 0.0;
      }
    }
  };
  int r = r_begin + 16 * wave;
  if (r < r_end) fetch(r, a, b);
  for (; r < r_end; r += 64) {
    if (r + 64 < r_end) fetch(r + 64, na, nb);
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int ti = 0; ti < TI; ++ti)
#pragma unroll
        for (int tj = 0; tj < TJ; ++tj)
          acc[ti][tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[ti][m], b[tj][m], acc[ti][tj], 0, 0, 0);
#pragma unroll
    for (int ti = 0; ti < TI; ++ti)
#pragma unroll
      for (int m = 0; m < 4; ++m) a[ti][m] = na[ti][m];
#pragma unroll
    for (int tj = 0; tj < TJ; ++tj)
#pragma unroll
      for (int m = 0; m < 4; ++m) b[tj][m] = nb[tj][m];
  }
  // value v of acc[ti][tj] in lane (lr, lk): D[16 ti + lk + 4 v][16 tj + lr]
#pragma unroll
  for (int ti = 0; ti < TI; ++ti)
#pragma unroll
    for (int tj = 0; tj < TJ; ++tj)
#pragma unroll
      for (int v = 0; v < 4; ++v) red[wave][16 * ti + lk + 4 * v][16 * tj + lr] = acc[ti][tj][v];
  __syncthreads();
  for (int s = tid; s < p * q; s += 256) {
    const int i = s / q, j = s - i * q;
    partial[(size_t)blockIdx.x * p * q + s] = ((red[0][i][j] + red[1][i][j]) + red[2][i][j]) + red[3][i][j];
  }
}

// out[v] = sum_b partial[b nvals + v] in a fixed order: a block takes 8
// values, lane group g (of 32) sums the partials b = g, g + 32, ... in
// order, then the 32 group sums are added in group order.  Neighbouring
// lanes read neighbouring values.  (Groups of 8 over 32 values left 64
// dependent adds per lane and few blocks: 18 us for 256 values of 512
// partials.)
constexpr int RV = 8, RG = 32;
__global__ void __launch_bounds__(256) k_gram_reduce(int nvals, int nb, const double* __restrict__ partial,
                                                     double* __restrict__ out) {
  __shared__ double red[RG][RV];
  const int vl = threadIdx.x % RV, g = threadIdx.x / RV;
  const int v = blockIdx.x * RV + vl;
  double s = 0.0;
  if (v < nvals)
    for (int b = g; b < nb; b += RG) s += partial[(size_t)b * nvals + v];
  red[g][vl] = s;
  __syncthreads();
  if (g == 0 && v < nvals) {
    double t = red[0][vl];
    for (int k = 1; k < RG; ++k) t += red[k][vl];
    out[v] = t;
  }
}

void launch_gram(int n, int p, int q, const double* U, long ldu, const double* V, long ldv, double* partial,
                 double* out, hipStream_t st) {
  if (p < 1 || q < 1) return;
  if (p > GP || q > GP) throw std::invalid_argument("Gram operands are limited to 96 columns");
  int nb = gram_blocks(n);
  int rpb = (n + nb - 1) / nb;
  rpb = (rpb + GR - 1) / GR * GR;
  const int w = p > q ? p : q;
  if (w <= 32) {
    const int ti = p > 16 ? 2 : 1, tj = q > 16 ? 2 : 1;
    if (ti == 1 && tj == 1)
      hipLaunchKernelGGL((k_gram_mfma<1, 1>), dim3(nb), dim3(256), 0, st, n, p, q, U, ldu, V, ldv, rpb, partial);
    else if (ti == 1)
      hipLaunchKernelGGL((k_gram_mfma<1, 2>), dim3(nb), dim3(256), 0, st, n, p, q, U, ldu, V, ldv, rpb, partial);
    else if (tj == 1)
      hipLaunchKernelGGL((k_gram_mfma<2, 1>), dim3(nb), dim3(256), 0, st, n, p, q, U, ldu, V, ldv, rpb, partial);
    else
      hipLaunchKernelGGL((k_gram_mfma<2, 2>), dim3(nb), dim3(256), 0, st, n, p, q, U, ldu, V, ldv, rpb, partial);
  } else if (w <= 64)
    hipLaunchKernelGGL(k_gram<4>, dim3(nb), dim3(256), 0, st, n, p, q, U, ldu, V, ldv, rpb, partial);
  else
    hipLaunchKernelGGL(k_gram<6>, dim3(nb), dim3(256), 0, st, n, p, q, U, ldu, V, ldv, rpb, partial);
  int nv = p * q;
  hipLaunchKernelGGL(k_gram_reduce, dim3((nv + RV - 1) / RV), dim3(256), 0, st, nv, nb, partial, out);
}

// Y[r][j] = alpha sum_c S[r][c] C[c][j] + beta Y[r][j], C row-major p x m
// (m <= 16 MT); Y is not read when beta == 0.  On the FP64 matrix cores as
// Y^T = C^T S^T, 64 rows of S at a time: the block stages the rows in LDS
// (column-major, coalesced loads; the next rows are loaded into registers
// meanwhile), wave w holds the C^T tile of output columns 16 (w % MT) .. + 15
// in registers for the whole launch (lane (lr, lk): C[4 k + lk][16 t + lr])
// and multiplies it into MT of the four 16-row tiles with
// v_mfma_f64_16x16x4f64 (value q of lane (lr, lk) is Y[r0 + lr][16 t + lk +
// 4 q]).  Output tiles from column jsplit on skip the first ksplit rows of C
// (zero there).  Every row of S a block reads is in LDS before any of the
// block's writes of that row, and blocks own disjoint rows, so Y may be
// columns of S itself (an in-place update).
constexpr int CS = 64;        // rows per stage
constexpr int CSP = CS + 16;  // padded LDS row: the four 16-lane groups of a read hit other banks
// KS: k steps of four columns of S (p <= 4 KS); BETA: Y is read
template <int MT, int KS, bool BETA>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) k_combine(int n, int p, int m, const double* S, long lds,
                                                 const double* __restrict__ C, double* Y, long ldy, double alpha,
                                                 double beta, int jsplit, int ksplit, int stages_per_block) {
  typedef double v4 __attribute__((ext_vector_type(4)));
  constexpr int WPT = 4 / MT;  // waves per output tile
  constexpr int RTW = 4 / WPT; // 16-row tiles per wave and stage
  // the stage's 4 KS columns only (p <= 4 KS): 20 KB at KS = 8 instead of
  // the 61 KB of the widest shape, so the LDS no longer holds every shape
  // to 2 waves per SIMD
  __shared__ double sS[4 * KS * CSP];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, lr = lane & 15, lk = lane >> 4;
  const int t = wave % MT, rw = wave / MT;
  const int pp = (p + 3) & ~3;
  const int kb = (16 * t >= jsplit) ? ksplit / 4 : 0;
  double a[KS];
#pragma unroll
  for (int k = 0; k < KS; ++k) {
    const int c = 4 * k + lk, j = 16 * t + lr;
    a[k] = (c < p && j < m) ? C[(size_t)c * m + j] : 0.0;
  }
  const long nst = ((long)n + CS - 1) / CS;
  const long s0 = (long)blockIdx.x * stages_per_block;
  const long s1 = s0 + stages_per_block < nst ? s0 + stages_per_block : nst;
  // staging: row, first column (then every 4th).  CS = 64: the column is
  // the wave index, uniform, so each load's column address is scalar
  static_assert(CS == 64, "one staged column per wave and load");
  const int lrow = lane, lcol = __builtin_amdgcn_readfirstlane(wave);
  double nx[KS];
  auto fetch = [&](long sg) {
    const long r = sg * CS + lrow;
#pragma unroll
    for (int u = 0; u < KS; ++u) {
      const int c = lcol + 4 * u;
      nx[u] = (c < p && r < n) ? S[(size_t)c * lds + r] : 0.0;
    }
  };
  if (s0 < s1) fetch(s0);
  for (long sg = s0; sg < s1; ++sg) {
    __syncthreads();  // the previous stage's LDS reads are done
#pragma unroll
    for (int u = 0; u < KS; ++u) {
      const int c = lcol + 4 * u;
      if (c < pp) sS[c * CSP + lrow] = nx[u];
    }
    __syncthreads();
    if (sg + 1 < s1) fetch(sg + 1);  // in flight during the products
    const long r0 = sg * CS;
    double yv[RTW][4];
    if (BETA) {
#pragma unroll
      for (int j = 0; j < RTW; ++j)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const long r = r0 + 16 * (rw * RTW + j) + lr;
          const int jj = 16 * t + lk + 4 * q;
          yv[j][q] = (jj < m && r < n) ? Y[(size_t)jj * ldy + r] : 0.0;
        }
    }
    v4 acc[RTW];
#pragma unroll
    for (int j = 0; j < RTW; ++j) acc[j] = v4{0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < KS; ++k) {
      if (k >= kb && 4 * k < pp) {
#pragma unroll
        for (int j = 0; j < RTW; ++j) {
          const double bv = sS[(4 * k + lk) * CSP + 16 * (rw * RTW + j) + lr];
          acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[k], bv, acc[j], 0, 0, 0);
        }
      }
      // keep the LDS reads of later k steps from being hoisted here (their
      // registers would cut the waves per SIMD)
      if (k % 4 == 3) __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int j = 0; j < RTW; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const long r = r0 + 16 * (rw * RTW + j) + lr;
        const int jj = 16 * t + lk + 4 * q;
        if (jj < m && r < n) {
          const double v = alpha * acc[j][q];
          Y[(size_t)jj * ldy + r] = BETA ? v + beta * yv[j][q] : v;
        }
      }
  }
}

static void combine_impl(int n, int p, int m, const double* S, long lds, const double* C, double* Y, long ldy,
                         double alpha, double beta, int jsplit, int ksplit, hipStream_t st) {
  if (p < 1 || m < 1) return;
  if (p > GP || m > CM) throw std::invalid_argument("block combination limited to 96 x 64 coefficients");
  const long nst = ((long)n + CS - 1) / CS;
  const int nb = (int)(nst < 1024 ? nst : 1024);
  const int spb = (int)((nst + nb - 1) / nb);
  const bool b = beta != 0.0;
  const int ks = p <= 32 ? 8 : p <= 64 ? 16 : 24;
  const int mt = m <= 16 ? 1 : m <= 32 ? 2 : 4;
#define TSEM_COMBINE(MT, KS, B)                                                                                   \
  if (mt == MT && ks == KS && b == B) {                                                                           \
    hipLaunchKernelGGL((k_combine<MT, KS, B>), dim3(nb), dim3(256), 0, st, n, p, m, S, lds, C, Y, ldy, alpha, beta, \
                       jsplit, ksplit, spb);                                                                      \
    return;                                                                                                       \
  }
  TSEM_COMBINE(1, 8, false) TSEM_COMBINE(2, 8, false) TSEM_COMBINE(4, 8, false)
  TSEM_COMBINE(1, 16, false) TSEM_COMBINE(2, 16, false) TSEM_COMBINE(4, 16, false)
  TSEM_COMBINE(1, 24, false) TSEM_COMBINE(2, 24, false) TSEM_COMBINE(4, 24, false)
  TSEM_COMBINE(1, 8, true) TSEM_COMBINE(2, 8, true) TSEM_COMBINE(4, 8, true)
  TSEM_COMBINE(1, 16, true) TSEM_COMBINE(2, 16, true) TSEM_COMBINE(4, 16, true)
  TSEM_COMBINE(1, 24, true) TSEM_COMBINE(2, 24, true) TSEM_COMBINE(4, 24, true)
#undef TSEM_COMBINE
  throw std::logic_error("block combination: no kernel for this shape");
}

void launch_combine(int n, int p, int m, const double* S, long lds, const double* C, double* Y, long ldy,
                    double alpha, double beta, hipStream_t st) {
  combine_impl(n, p, m, S, lds, C, Y, ldy, alpha, beta, m, 0, st);
}

void launch_combine_split(int n, int p, int m, const double* S, long lds, const double* C, double* Y, long ldy,
                          int jsplit, int ksplit, hipStream_t st) {
  if (jsplit % 16 || ksplit % 4 || ksplit > p) throw std::invalid_argument("combination split off its tiles");
  combine_impl(n, p, m, S, lds, C, Y, ldy, 1.0, 0.0, jsplit, ksplit, st);
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
