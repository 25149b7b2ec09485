// Element operator on the matrix cores (gfx950 MFMA).
//
// k_gather_tiles: element-local coefficients of 32-element tiles, written
//   k-major inside the tile (xt[tile][k][32]) so the GEMM kernel streams them
//   with coalesced loads.
// k_elem_gemm: for a block of NE = 16 CT elements
//   U (3 Mp x NE) = D (3 Mp x K) X (K x NE)            D_a: derivative factors
//   V_a = sum_b Ghat_ab(e) U_b                          per element column
//   Y (K x NE)   = D^T (K x 3 Mp) V (3 Mp x NE)
//   with v_mfma_{f32,f64}_16x16x4.  Each a-block of rows is padded to Mp (a
//   multiple of 16) so the three derivative components of one (mode,
//   element) sit in the same lane and the 3x3 mixing happens in registers;
//   V goes through LDS between the GEMMs.  D is block triangular in degree
//   (d_a phi_i has degree deg(phi_i) - 1): per tile the k ranges skip the
//   structurally zero blocks (about 30% of the dense work at p = 6).
//   Y leaves through an LDS transpose, element-major (yl[e][i]), so the
//   scatter reads each element's modes from contiguous memory.
// k_mass_gemm: the mass term, Y = M X with the reference mass matrix, each
//   element column scaled by its volume factor; stored, or added to the
//   stiffness kernel's output.
//
// MFMA 16x16x4 operand layout (one 64-lane wave): lane l supplies
// A[l % 16][l / 16] and B[l / 16][l % 16]; accumulator value j of lane l is
// C[4 (l / 16) + j][l % 16] for f32 and C[l / 16 + 4 j][l % 16] for f64
// (measured on gfx950).
#include <hip/hip_runtime.h>

#include <stdexcept>

#include "kernels.hpp"

namespace tsem {

namespace {
constexpr int WAVES = 4;
constexpr int NT = 64 * WAVES;
constexpr int RMAX = 12;  // row tiles of GEMM 2 per block (K <= 192)

template <typename T> struct Mma;
template <> struct Mma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ inline v4 op(float a, float b, v4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  static __device__ inline int row(int lk, int j) { return 4 * lk + j; }
};
template <> struct Mma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ inline v4 op(double a, double b, v4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
  static __device__ inline int row(int lk, int j) { return lk + 4 * j; }
};

template <typename T>
__device__ inline T ldx(const T* __restrict__ x, const uint8_t* __restrict__ cmask, int g, int constrained) {
  if (constrained && cmask[g]) return T(0);
  return x[g];
}

template <typename T>
__device__ inline T gather_one(const DevSpace& S, int e, int i, const T* __restrict__ x, int constrained) {
  const ElemDesc& D = S.desc[e];
  int code = S.slot[i];
  int kind = code & 3, geo = (code >> 2) & 7, sub = code >> 5;
  if (kind == 0) return ldx(x, S.cmask, D.v[geo], constrained);
  if (kind == 1) {
    int ed = D.e[geo];
    T v = ldx(x, S.cmask, (ed & 0x7fffffff) + sub, constrained);
    return (ed < 0 && (sub & 1)) ? -v : v;
  }
  if (kind == 3) return ldx(x, S.cmask, D.t + sub, constrained);
  int fo = D.fo[geo];
  int row = (fo & 7) * S.nface + sub;
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
    acc += ldx(x, S.cmask, g, constrained) * w;
  }
  return acc;
}

}  // namespace

// one block per tile of TILE elements: the tile's element descriptors are
// staged in LDS once, then each thread produces slots (k, element) for
// k < Kl rounded up to 4; writes zeros for k >= Kl and for padding elements.  Dirichlet-eliminated DOFs
// read as zero through the per-element geometry bits (vertex k: bit k, edge
// k: bit 4+k, face k: bit 10+k, interior: bit 14).
template <typename T>
__global__ void __launch_bounds__(256) k_gather_tiles(DevSpace S, int Kl, const T* __restrict__ x, T* __restrict__ xt,
                                                      int constrained) {
  __shared__ ElemDesc sd[TILE];
  __shared__ int sm[TILE];
  const int tile = xcd_block(blockIdx.x, gridDim.x);
  const int e0 = tile * TILE;
  // the face table rows of this space: w [R J], n [R], key [R J], then the slot codes [K]
  extern __shared__ unsigned char gft_raw[];
  const int R = 6 * S.ndg2, J = S.ftc_J;
  T* cw = reinterpret_cast<T*>(gft_raw);
  int* cn = reinterpret_cast<int*>(cw + (size_t)R * J);
  int* ck = cn + R;
  int* sslot = ck + R * J;  // the slot codes of the local modes: no global load per entry
  // faces in a diagonal variant (bit v of sl_diag: each local face mode is
  // +- the global face DOF of the same index, sign bit d of sl_sgn[v]) skip
  // the table rows: one load, the same value as the one-term row gives
  __shared__ uint32_t ssgn[6];
  const int dg = S.sl_diag;
  // the set-up loads are all issued before any of them is waited for (as in
  // k_gather_tiles_h); table entries past the first TQ per thread follow in
  // plain loops
  {
    const int tid = threadIdx.x;
    const int nvalid = min(TILE, S.nt - e0);
    constexpr int W = sizeof(ElemDesc) / 4;
    constexpr int DQ = (TILE * W + 255) / 256;
    constexpr int TQ = 2;
    const int32_t* src = reinterpret_cast<const int32_t*>(S.desc + e0);
    int32_t* dst = reinterpret_cast<int32_t*>(sd);
    int32_t dv[DQ];
#pragma unroll
    for (int q = 0; q < DQ; ++q) {
      const int s = tid + 256 * q;
      dv[q] = (s < TILE * W && s / W < nvalid) ? src[s] : 0;
    }
    const int gmv = (constrained && tid < nvalid) ? S.gmask[e0 + tid] : 0;
    T cwv[TQ];
    int ckv[TQ];
#pragma unroll
    for (int q = 0; q < TQ; ++q) {
      const int s = tid + 256 * q;
      cwv[q] = s < R * J ? ((sizeof(T) == 8) ? (T)S.ftc_w[s] : (T)S.ftc_w32[s]) : T(0);
      ckv[q] = s < R * J ? S.ftc_key[s] : 0;
    }
    const int cnv = tid < R ? S.ftc_n[tid] : 0;
    const int slv = tid < S.K ? S.slot[tid] : 0;
    const uint32_t sgv = tid < 6 ? S.sl_sgn[tid] : 0u;
#pragma unroll
    for (int q = 0; q < DQ; ++q) {
      const int s = tid + 256 * q;
      if (s < TILE * W) dst[s] = dv[q];
    }
    if (tid < TILE) sm[tid] = gmv;
#pragma unroll
    for (int q = 0; q < TQ; ++q) {
      const int s = tid + 256 * q;
      if (s < R * J) { cw[s] = cwv[q]; ck[s] = ckv[q]; }
    }
    for (int s = tid + 256 * TQ; s < R * J; s += 256) {
      cw[s] = (sizeof(T) == 8) ? (T)S.ftc_w[s] : (T)S.ftc_w32[s];
      ck[s] = S.ftc_key[s];
    }
    if (tid < R) cn[tid] = cnv;
    for (int s = tid + 256; s < R; s += 256) cn[s] = S.ftc_n[s];
    if (tid < S.K) sslot[tid] = slv;
    for (int s = tid + 256; s < S.K; s += 256) sslot[s] = S.slot[s];
    if (tid < 6) ssgn[tid] = sgv;
  }
  __syncthreads();
  const int Kp = S.Kp;
  const int Kpl = (Kl + 3) & ~3;  // the GEMM reads rows k < Kpl only
  T* out = xt + (size_t)tile * Kp * TILE;
  for (int r = threadIdx.x; r < Kpl * TILE; r += 256) {
    int k = r / TILE, le = r % TILE, e = e0 + le;
    T v = T(0);
    if (e < S.nt && k < Kl) {
      const ElemDesc& D = sd[le];
      const int msk = sm[le];
      int code = sslot[k];
      int kind = code & 3, geo = (code >> 2) & 7, sub = code >> 5;
      if (kind == 0) {
        v = ((msk >> geo) & 1) ? T(0) : x[D.v[geo]];
      } else if (kind == 1) {
        int ed = D.e[geo];
        v = ((msk >> (4 + geo)) & 1) ? T(0) : x[(ed & 0x7fffffff) + sub];
        if (ed < 0 && (sub & 1)) v = -v;
      } else if (kind == 3) {
        v = ((msk >> 14) & 1) ? T(0) : x[D.t + sub];
      } else {
        int fo = D.fo[geo];
        const bool fmask = (msk >> (10 + geo)) & 1;
        if ((dg >> (fo & 7)) & 1) {
          v = fmask ? T(0) : x[D.f[geo] + sub];
          if (sub < 32 && ((ssgn[fo & 7] >> sub) & 1)) v = -v;
        } else {
          const int row = (fo & 7) * S.ndg2 + sub;
          T acc = T(0);
          const int nterm = cn[row];
          for (int j = 0; j < nterm; ++j) {
            int key = ck[row * J + j];
            int ie = key & 3, idx = key >> 2;
            T w = cw[row * J + j];
            T xv;
            if (ie < 3) {
              int le2 = (fo >> (3 + 3 * ie)) & 7;
              if (((fo >> (12 + ie)) & 1) && (idx & 1)) w = -w;
              xv = ((msk >> (4 + le2)) & 1) ? T(0) : x[(D.e[le2] & 0x7fffffff) + idx];
            } else {
              xv = fmask ? T(0) : x[D.f[geo] + idx];
            }
            acc += xv * w;
          }
          v = acc;
        }
      }
    }
    out[r] = v;
  }
}

template <typename T, int CT>
__global__ void __launch_bounds__(NT) k_elem_gemm(DevSpace S, OpCoef op, const T* __restrict__ Dp,
                                                  const T* __restrict__ xt, T* __restrict__ yl) {
  using v4 = typename Mma<T>::v4;
  constexpr int NE = 16 * CT;
  constexpr int VS = NE + 1;  // odd row stride of the LDS tiles: no bank conflicts
  extern __shared__ unsigned char smem_raw[];
  T* V = reinterpret_cast<T*>(smem_raw);  // [3 Mp][VS]
  const int K = S.K;
  const int Kl = op.Kl > 0 ? op.Kl : K;
  const int Kdl = op.Kdl > 0 ? op.Kdl : S.Kd;
  const int Mp = S.Mp, Cp = S.Cp, Kp = S.Kp;
  const int Mt = (Kdl + 15) >> 4;
  const int Rt = (Kl + 15) >> 4;
  const int Kpl = (Kl + 3) & ~3;
  // wave-uniform in a scalar register: the per-wave tables (GEMM row starts
  // and ends) are read with scalar loads, with no vector-memory wait
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lr = lane & 15, lk = lane >> 4;
  const int e0 = blockIdx.x * NE;
  const T cst = (T)op.stiff;

  // ---- GEMM 1 + mixing: wave w owns m tiles w, w + WAVES, ...
  for (int t = wave; t < Mt; t += WAVES) {
    v4 acc[3][CT];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int c = 0; c < CT; ++c) acc[a][c] = v4{0, 0, 0, 0};
    const T* d0 = Dp + (size_t)(0 * Mp + 16 * t + lr) * Cp + lk;
    const T* d1 = Dp + (size_t)(1 * Mp + 16 * t + lr) * Cp + lk;
    const T* d2 = Dp + (size_t)(2 * Mp + 16 * t + lr) * Cp + lk;
    const T* xb[CT];
#pragma unroll
    for (int c = 0; c < CT; ++c) {
      int le = 16 * c + lr, e = e0 + le;
      xb[c] = xt + (size_t)(e >> 5) * Kp * TILE + (size_t)lk * TILE + (e & 31);
    }
    const int kb = S.kbeg1[t];
#pragma unroll 2
    for (int k = kb; k < Kpl; k += 4) {
      T a0 = d0[k], a1 = d1[k], a2 = d2[k];
#pragma unroll
      for (int c = 0; c < CT; ++c) {
        T b = xb[c][(size_t)k * TILE];
        acc[0][c] = Mma<T>::op(a0, b, acc[0][c]);
        acc[1][c] = Mma<T>::op(a1, b, acc[1][c]);
        acc[2][c] = Mma<T>::op(a2, b, acc[2][c]);
      }
    }
#pragma unroll
    for (int c = 0; c < CT; ++c) {
      int col = 16 * c + lr, e = e0 + col;
      T g0 = 0, g1 = 0, g2 = 0, g3 = 0, g4 = 0, g5 = 0;
      if (e < S.nt) {
        const T* g = (sizeof(T) == 8) ? (const T*)(S.ghat + 6 * (size_t)e) : (const T*)(S.ghat32 + 6 * (size_t)e);
        g0 = cst * g[0]; g1 = cst * g[1]; g2 = cst * g[2]; g3 = cst * g[3]; g4 = cst * g[4]; g5 = cst * g[5];
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        int m = 16 * t + Mma<T>::row(lk, j);
        T u0 = acc[0][c][j], u1 = acc[1][c][j], u2 = acc[2][c][j];
        T keep = (m < Kdl) ? T(1) : T(0);
        V[(size_t)(0 * Mp + m) * VS + col] = keep * (g0 * u0 + g3 * u1 + g4 * u2);
        V[(size_t)(1 * Mp + m) * VS + col] = keep * (g3 * u0 + g1 * u1 + g5 * u2);
        V[(size_t)(2 * Mp + m) * VS + col] = keep * (g4 * u0 + g5 * u1 + g2 * u2);
      }
    }
  }
  __syncthreads();

  // ---- GEMM 2: Y[i][e] = sum_r Dp[r][i] V[r][e].  Wave w owns row tiles
  // w + WAVES j and every column tile; per 16 rows it issues all factor
  // loads (global) and V loads (LDS) first, then the MFMAs, so each factor
  // value serves all CT column tiles
  constexpr int RJ = (RMAX + WAVES - 1) / WAVES;
  v4 acc[RJ][CT];
#pragma unroll
  for (int j = 0; j < RJ; ++j)
#pragma unroll
    for (int c = 0; c < CT; ++c) acc[j][c] = v4{0, 0, 0, 0};
  int kend[RJ];
  int kmax = 0;
#pragma unroll
  for (int j = 0; j < RJ; ++j) {
    int t2 = wave + WAVES * j;
    kend[j] = (t2 < Rt) ? min(S.kend2[t2], 16 * Mt) : 0;
    kmax = max(kmax, kend[j]);
  }
  for (int a = 0; a < 3; ++a) {
    for (int k0 = 0; k0 < kmax; k0 += 16) {
      T av[4][RJ], bv[4][CT];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int k = k0 + 4 * u;
#pragma unroll
        for (int j = 0; j < RJ; ++j) {
          int t2 = wave + WAVES * j;
          int i = 16 * (t2 < Rt ? t2 : 0) + lr;
          av[u][j] = (k < kmax) ? Dp[(size_t)(a * Mp + k + lk) * Cp + i] : T(0);
        }
#pragma unroll
        for (int c = 0; c < CT; ++c) bv[u][c] = (k < kmax) ? V[(size_t)(a * Mp + k + lk) * VS + 16 * c + lr] : T(0);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int k = k0 + 4 * u;
#pragma unroll
        for (int j = 0; j < RJ; ++j)
          if (k < kend[j]) {
#pragma unroll
            for (int c = 0; c < CT; ++c) acc[j][c] = Mma<T>::op(av[u][j], bv[u][c], acc[j][c]);
          }
      }
    }
  }
  // ---- write Y element-major (yl[e][i]) through an LDS transpose
  __syncthreads();  // every wave is done reading V
  T* Y = V;         // [K][VS]
#pragma unroll
  for (int j = 0; j < RJ; ++j) {
    int t2 = wave + WAVES * j;
    if (t2 < Rt) {
#pragma unroll
      for (int c = 0; c < CT; ++c)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          int i = 16 * t2 + Mma<T>::row(lk, q);
          if (i < K) Y[(size_t)i * VS + 16 * c + lr] = acc[j][c][q];
        }
    }
  }
  __syncthreads();
  const int nval = min(NE, S.nt - e0);
  for (int s = tid; s < nval * K; s += NT) {
    int le = s / K, i = s % K;
    yl[(size_t)(e0 + le) * K + i] = (i < Kl && i < 16 * Rt) ? Y[(size_t)i * VS + le] : T(0);
  }
}

// The stiffness operator with the gather fused in: the block's element
// coefficients go from x straight into an LDS tile X[k][element] (the layout
// the tiled gather writes to HBM, here never leaving the CU), and GEMM 1 reads
// its B operand from there.  X and the staging tables share the LDS of the V
// tile: each wave keeps its V values in registers until every wave is done
// with X, then writes them.  The gather is that of k_gather_tiles (descriptors,
// masks, face-table rows and slot codes staged in LDS, all staging loads
// issued before the first wait); the GEMMs and the Y write are those of
// k_elem_gemm.  Waves per SIMD: 3 in FP64 (194 registers before the bound),
// 4 in FP32 (134 registers before the bound, 128 with a few spills: on the
// 32^3 mesh 55.3 -> 47.6 us at p = 3 and 88.1 -> 77.7 us at p = 4 against 3).
template <typename T, int CT>
__global__ void __launch_bounds__(NT, sizeof(T) == 4 ? 4 : 3) k_elem_gemm_gx(DevSpace S, OpCoef op, const T* __restrict__ Dp,
                                                     const T* __restrict__ x, T* __restrict__ yl) {
  using v4 = typename Mma<T>::v4;
  constexpr int NE = 16 * CT;
  constexpr int VS = NE + 1;
  constexpr int TPW = 2;  // m tiles per wave: Mt <= 8 (Mp <= 128)
  extern __shared__ unsigned char smem_raw[];
  T* V = reinterpret_cast<T*>(smem_raw);  // [3 Mp][VS], after GEMM 1
  const int K = S.K;
  const int Kl = op.Kl > 0 ? op.Kl : K;
  const int Kdl = op.Kdl > 0 ? op.Kdl : S.Kd;
  const int Mp = S.Mp, Cp = S.Cp, Kp = S.Kp;
  const int Mt = (Kdl + 15) >> 4;
  const int Rt = (Kl + 15) >> 4;
  const int Kpl = (Kl + 3) & ~3;
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lr = lane & 15, lk = lane >> 4;
  const int e0 = xcd_block(blockIdx.x, gridDim.x) * NE;  // contiguous element range per XCD
  const int nvalid = min(NE, S.nt - e0);
  const T cst = (T)op.stiff;
  x += (size_t)blockIdx.y * op.ldx;  // column of a block apply
  yl += (size_t)blockIdx.y * op.ldl;
  // ---- phase 0: staging and the gather into X
  const int R = 6 * S.ndg2, J = S.ftc_J;
  T* X = reinterpret_cast<T*>(smem_raw);  // [Kp][VS]
  ElemDesc* sd = reinterpret_cast<ElemDesc*>(X + (size_t)Kp * VS);
  int* sm = reinterpret_cast<int*>(sd + NE);
  T* cw = reinterpret_cast<T*>(sm + NE + (NE & 1));  // 8-byte aligned
  int* cn = reinterpret_cast<int*>(cw + (size_t)R * J);
  int* ck = cn + R;
  int* sslot = ck + (size_t)R * J;
  uint32_t* ssgn = reinterpret_cast<uint32_t*>(sslot + K);
  const int dg = S.sl_diag;
  {
    constexpr int W = sizeof(ElemDesc) / 4;
    constexpr int DQ = (NE * W + NT - 1) / NT;
    constexpr int TQ = 2;
    const int32_t* src = reinterpret_cast<const int32_t*>(S.desc + e0);
    int32_t dv[DQ];
#pragma unroll
    for (int q = 0; q < DQ; ++q) {
      const int s2 = tid + NT * q;
      dv[q] = (s2 < NE * W && s2 / W < nvalid) ? src[s2] : 0;
    }
    const int gmv = (op.constrained && tid < nvalid) ? S.gmask[e0 + tid] : 0;
    T cwv[TQ];
    int ckv[TQ];
#pragma unroll
    for (int q = 0; q < TQ; ++q) {
      const int s2 = tid + NT * q;
      cwv[q] = s2 < R * J ? ((sizeof(T) == 8) ? (T)S.ftc_w[s2] : (T)S.ftc_w32[s2]) : T(0);
      ckv[q] = s2 < R * J ? S.ftc_key[s2] : 0;
    }
    const int cnv = tid < R ? S.ftc_n[tid] : 0;
    const int slv = tid < K ? S.slot[tid] : 0;
    const uint32_t sgv = tid < 6 ? S.sl_sgn[tid] : 0u;
    int32_t* dst = reinterpret_cast<int32_t*>(sd);
#pragma unroll
    for (int q = 0; q < DQ; ++q) {
      const int s2 = tid + NT * q;
      if (s2 < NE * W) dst[s2] = dv[q];
    }
    if (tid < NE) sm[tid] = gmv;
#pragma unroll
    for (int q = 0; q < TQ; ++q) {
      const int s2 = tid + NT * q;
      if (s2 < R * J) { cw[s2] = cwv[q]; ck[s2] = ckv[q]; }
    }
    for (int s2 = tid + NT * TQ; s2 < R * J; s2 += NT) {
      cw[s2] = (sizeof(T) == 8) ? (T)S.ftc_w[s2] : (T)S.ftc_w32[s2];
      ck[s2] = S.ftc_key[s2];
    }
    if (tid < R) cn[tid] = cnv;
    for (int s2 = tid + NT; s2 < R; s2 += NT) cn[s2] = S.ftc_n[s2];
    if (tid < K) sslot[tid] = slv;
    for (int s2 = tid + NT; s2 < K; s2 += NT) sslot[s2] = S.slot[s2];
    if (tid < 6) ssgn[tid] = sgv;
  }
  __syncthreads();
  for (int r = tid; r < Kpl * NE; r += NT) {
    const int k = r / NE, le = r % NE;
    T v = T(0);
    if (le < nvalid && k < Kl) {
      const ElemDesc& D = sd[le];
      const int msk = sm[le];
      const int code = sslot[k];
      const int kind = code & 3, geo = (code >> 2) & 7, sub = code >> 5;
      if (kind == 0) {
        v = ((msk >> geo) & 1) ? T(0) : x[D.v[geo]];
      } else if (kind == 1) {
        const int ed = D.e[geo];
        v = ((msk >> (4 + geo)) & 1) ? T(0) : x[(ed & 0x7fffffff) + sub];
        if (ed < 0 && (sub & 1)) v = -v;
      } else if (kind == 3) {
        v = ((msk >> 14) & 1) ? T(0) : x[D.t + sub];
      } else {
        const int fo = D.fo[geo];
        const bool fmask = (msk >> (10 + geo)) & 1;
        if ((dg >> (fo & 7)) & 1) {
          v = fmask ? T(0) : x[D.f[geo] + sub];
          if (sub < 32 && ((ssgn[fo & 7] >> sub) & 1)) v = -v;
        } else {
          const int row = (fo & 7) * S.ndg2 + sub;
          T acc = T(0);
          const int nterm = cn[row];
          for (int j = 0; j < nterm; ++j) {
            const int key = ck[row * J + j];
            const int ie = key & 3, idx = key >> 2;
            T w = cw[row * J + j];
            T xq;
            if (ie < 3) {
              const int le2 = (fo >> (3 + 3 * ie)) & 7;
              if (((fo >> (12 + ie)) & 1) && (idx & 1)) w = -w;
              xq = ((msk >> (4 + le2)) & 1) ? T(0) : x[(D.e[le2] & 0x7fffffff) + idx];
            } else {
              xq = fmask ? T(0) : x[D.f[geo] + idx];
            }
            acc += xq * w;
          }
          v = acc;
        }
      }
    }
    X[(size_t)k * VS + le] = v;
  }
  __syncthreads();

  // ---- GEMM 1 + mixing into registers: wave w owns m tiles w, w + WAVES
  T vv[TPW][CT][3][4];
#pragma unroll
  for (int tp = 0; tp < TPW; ++tp) {
    const int t = wave + WAVES * tp;
    if (t >= Mt) break;
    v4 acc[3][CT];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int c = 0; c < CT; ++c) acc[a][c] = v4{0, 0, 0, 0};
    const T* d0 = Dp + (size_t)(0 * Mp + 16 * t + lr) * Cp + lk;
    const T* d1 = Dp + (size_t)(1 * Mp + 16 * t + lr) * Cp + lk;
    const T* d2 = Dp + (size_t)(2 * Mp + 16 * t + lr) * Cp + lk;
    const int kb = S.kbeg1[t];
#pragma unroll 2
    for (int k = kb; k < Kpl; k += 4) {
      const T a0 = d0[k], a1 = d1[k], a2 = d2[k];
#pragma unroll
      for (int c = 0; c < CT; ++c) {
        const T b = X[(size_t)(k + lk) * VS + 16 * c + lr];
        acc[0][c] = Mma<T>::op(a0, b, acc[0][c]);
        acc[1][c] = Mma<T>::op(a1, b, acc[1][c]);
        acc[2][c] = Mma<T>::op(a2, b, acc[2][c]);
      }
    }
#pragma unroll
    for (int c = 0; c < CT; ++c) {
      const int col = 16 * c + lr, e = e0 + col;
      T g0 = 0, g1 = 0, g2 = 0, g3 = 0, g4 = 0, g5 = 0;
      if (e < S.nt) {
        const T* g = (sizeof(T) == 8) ? (const T*)(S.ghat + 6 * (size_t)e) : (const T*)(S.ghat32 + 6 * (size_t)e);
        g0 = cst * g[0]; g1 = cst * g[1]; g2 = cst * g[2]; g3 = cst * g[3]; g4 = cst * g[4]; g5 = cst * g[5];
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int m = 16 * t + Mma<T>::row(lk, j);
        const T u0 = acc[0][c][j], u1 = acc[1][c][j], u2 = acc[2][c][j];
        const T keep = (m < Kdl) ? T(1) : T(0);
        vv[tp][c][0][j] = keep * (g0 * u0 + g3 * u1 + g4 * u2);
        vv[tp][c][1][j] = keep * (g3 * u0 + g1 * u1 + g5 * u2);
        vv[tp][c][2][j] = keep * (g4 * u0 + g5 * u1 + g2 * u2);
      }
    }
  }
  __syncthreads();  // every wave is done reading X: V may take its place
#pragma unroll
  for (int tp = 0; tp < TPW; ++tp) {
    const int t = wave + WAVES * tp;
    if (t >= Mt) break;
#pragma unroll
    for (int c = 0; c < CT; ++c)
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          V[(size_t)(a * Mp + 16 * t + Mma<T>::row(lk, j)) * VS + 16 * c + lr] = vv[tp][c][a][j];
  }
  __syncthreads();

  // ---- GEMM 2 and the Y write, as in k_elem_gemm
  constexpr int RJ = (RMAX + WAVES - 1) / WAVES;
  v4 acc[RJ][CT];
#pragma unroll
  for (int j = 0; j < RJ; ++j)
#pragma unroll
    for (int c = 0; c < CT; ++c) acc[j][c] = v4{0, 0, 0, 0};
  int kend[RJ];
  int kmax = 0;
#pragma unroll
  for (int j = 0; j < RJ; ++j) {
    const int t2 = wave + WAVES * j;
    kend[j] = (t2 < Rt) ? min(S.kend2[t2], 16 * Mt) : 0;
    kmax = max(kmax, kend[j]);
  }
  for (int a = 0; a < 3; ++a) {
    for (int k0 = 0; k0 < kmax; k0 += 16) {
      T av[4][RJ], bv[4][CT];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int k = k0 + 4 * u;
#pragma unroll
        for (int j = 0; j < RJ; ++j) {
          const int t2 = wave + WAVES * j;
          const int i = 16 * (t2 < Rt ? t2 : 0) + lr;
          av[u][j] = (k < kmax) ? Dp[(size_t)(a * Mp + k + lk) * Cp + i] : T(0);
        }
#pragma unroll
        for (int c = 0; c < CT; ++c) bv[u][c] = (k < kmax) ? V[(size_t)(a * Mp + k + lk) * VS + 16 * c + lr] : T(0);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int k = k0 + 4 * u;
#pragma unroll
        for (int j = 0; j < RJ; ++j)
          if (k < kend[j]) {
#pragma unroll
            for (int c = 0; c < CT; ++c) acc[j][c] = Mma<T>::op(av[u][j], bv[u][c], acc[j][c]);
          }
      }
    }
  }
  __syncthreads();  // every wave is done reading V
  T* Y = V;         // [K][VS]
#pragma unroll
  for (int j = 0; j < RJ; ++j) {
    const int t2 = wave + WAVES * j;
    if (t2 < Rt) {
#pragma unroll
      for (int c = 0; c < CT; ++c)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int i = 16 * t2 + Mma<T>::row(lk, q);
          if (i < K) Y[(size_t)i * VS + 16 * c + lr] = acc[j][c][q];
        }
    }
  }
  __syncthreads();
  for (int s2 = tid; s2 < nvalid * K; s2 += NT) {
    const int le = s2 / K, i = s2 % K;
    yl[(size_t)(e0 + le) * K + i] = (i < Kl && i < 16 * Rt) ? Y[(size_t)i * VS + le] : T(0);
  }
}

// Mass term on the matrix cores: Y (K x NE) = M (K x K) X (K x NE) with the
// reference mass matrix M = mfac^T mfac (Cp x Cp, zero padded), column e
// scaled by mass 6|V_e|; the same operand layout as GEMM 1 above.  Wave w
// owns the row tiles w + WAVES j.  With `accumulate` the result is added to
// yl (the stiffness kernel has written it), else it is stored.
template <typename T, int CT>
__global__ void __launch_bounds__(NT) k_mass_gemm(DevSpace S, OpCoef op, const T* __restrict__ Mm,
                                                  const T* __restrict__ xt, T* __restrict__ yl, int accumulate) {
  using v4 = typename Mma<T>::v4;
  constexpr int NE = 16 * CT;
  constexpr int VS = NE + 1;
  constexpr int RJ = (RMAX + WAVES - 1) / WAVES;
  extern __shared__ unsigned char smem_raw[];
  T* Y = reinterpret_cast<T*>(smem_raw);  // [Cp][VS]
  const int K = S.K;
  const int Kl = op.Kl > 0 ? op.Kl : K;
  const int Cp = S.Cp, Kp = S.Kp;
  const int Rt = (Kl + 15) >> 4;
  const int Kpl = (Kl + 3) & ~3;
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lr = lane & 15, lk = lane >> 4;
  const int e0 = blockIdx.x * NE;
  const T* xb[CT];
  T scl[CT];
#pragma unroll
  for (int c = 0; c < CT; ++c) {
    const int e = e0 + 16 * c + lr;
    xb[c] = xt + (size_t)(e >> 5) * Kp * TILE + (size_t)lk * TILE + (e & 31);
    scl[c] = e < S.nt ? (T)op.mass * ((sizeof(T) == 8) ? (T)S.sixv[e] : (T)S.sixv32[e]) : T(0);
  }
  v4 acc[RJ][CT];
#pragma unroll
  for (int j = 0; j < RJ; ++j)
#pragma unroll
    for (int c = 0; c < CT; ++c) acc[j][c] = v4{0, 0, 0, 0};
  for (int k = 0; k < Kpl; k += 4) {
    T b[CT];
#pragma unroll
    for (int c = 0; c < CT; ++c) b[c] = xb[c][(size_t)k * TILE];
#pragma unroll
    for (int j = 0; j < RJ; ++j) {
      const int t = wave + WAVES * j;
      if (t < Rt) {
        const T a = Mm[(size_t)(16 * t + lr) * Cp + k + lk];
#pragma unroll
        for (int c = 0; c < CT; ++c) acc[j][c] = Mma<T>::op(a, b[c], acc[j][c]);
      }
    }
  }
#pragma unroll
  for (int j = 0; j < RJ; ++j) {
    const int t = wave + WAVES * j;
    if (t < Rt) {
#pragma unroll
      for (int c = 0; c < CT; ++c)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int i = 16 * t + Mma<T>::row(lk, q);
          Y[(size_t)i * VS + 16 * c + lr] = acc[j][c][q] * scl[c];
        }
    }
  }
  __syncthreads();
  const int nval = min(NE, S.nt - e0);
  for (int s = tid; s < nval * K; s += NT) {
    const int le = s / K, i = s % K;
    const T v = i < Kl ? Y[(size_t)i * VS + le] : T(0);
    T* o = yl + (size_t)(e0 + le) * K + i;
    *o = accumulate ? *o + v : v;
  }
}

template <typename T, int CT>
static size_t mass_lds(const DevSpace& S) {
  return (size_t)S.Cp * (16 * CT + 1) * sizeof(T);
}

template <typename T, int CT>
static size_t gemm_lds(const DevSpace& S) {
  size_t rows = (size_t)3 * S.Mp > (size_t)S.K ? (size_t)3 * S.Mp : (size_t)S.K;
  return rows * (16 * CT + 1) * sizeof(T);
}

// LDS of the fused kernel: the V / Y tile, or the X tile with the staging
// tables, whichever is larger
template <typename T, int CT>
static size_t gemm_gx_lds(const DevSpace& S) {
  const size_t NE = 16 * CT, VS = NE + 1, R = 6 * (size_t)S.ndg2, J = S.ftc_J;
  const size_t stage = (size_t)S.Kp * VS * sizeof(T) + NE * sizeof(ElemDesc) + (NE + (NE & 1)) * 4 +
                       R * J * (sizeof(T) + 4) + R * 4 + (size_t)S.K * 4 + 24;
  const size_t v = gemm_lds<T, CT>(S);
  return v > stage ? v : stage;
}

bool mfma_supported(const DevSpace& S) {
  return S.Dp64 != nullptr && S.Mm64 != nullptr && S.xt64 != nullptr && S.K <= 16 * RMAX &&
         gemm_lds<float, 4>(S) <= 160 * 1024 && gemm_lds<double, 2>(S) <= 160 * 1024 &&
         mass_lds<float, 4>(S) <= 160 * 1024 && mass_lds<double, 2>(S) <= 160 * 1024;
}

bool mfma_gx_supported(const DevSpace& S) {
  return mfma_supported(S) && S.Mp <= 128 && gemm_gx_lds<float, 4>(S) <= 160 * 1024 &&
         gemm_gx_lds<double, 2>(S) <= 160 * 1024;
}

// stiffness (k_elem_gemm) and mass (k_mass_gemm) terms from one gather; an
// operator with neither runs the stiffness kernel, which then writes zeros
template <typename T, int CT>
static void run(const DevSpace& S, const OpCoef& op, const T* Dp, const T* Mm, T* xt, const T* x, T* yl,
                hipStream_t st) {
  const int Kl = op.Kl > 0 ? op.Kl : S.K;
  const bool timed = op.timed();
  const int kind = sizeof(T) == 8 ? TK_GATHER64 : TK_GATHER32;
  int nb = (S.nt + 16 * CT - 1) / (16 * CT);
  const bool mass = op.mass != 0.0, stiff = op.stiff != 0.0 || !mass;
  if (stiff && !mass && mfma_gx_supported(S)) {
    // stiffness alone: the gather fused into the element kernel (timed as
    // it), over op.ncol columns
    const int th1 = timed ? timing_begin(kind + 1, st) : -1;
    hipLaunchKernelGGL((k_elem_gemm_gx<T, CT>), dim3(nb, op.ncol), dim3(NT), (gemm_gx_lds<T, CT>(S)), st, S, op, Dp,
                       x, yl);
    timing_end(th1, st);
    return;
  }
  if (op.ncol != 1) throw std::logic_error("a block of columns needs the stiffness operator alone");
  int th = timed ? timing_begin(kind, st) : -1;
  const size_t ftc = (size_t)6 * S.ndg2 * (S.ftc_J * (sizeof(T) + 4) + 4) + (size_t)S.K * 4;
  hipLaunchKernelGGL(k_gather_tiles<T>, dim3((unsigned)S.ntile), dim3(256), ftc, st, S, Kl, x, xt, op.constrained);
  timing_end(th, st);
  th = timed ? timing_begin(kind + 1, st) : -1;
  if (stiff) hipLaunchKernelGGL((k_elem_gemm<T, CT>), dim3(nb), dim3(NT), (gemm_lds<T, CT>(S)), st, S, op, Dp, xt, yl);
  if (mass)
    hipLaunchKernelGGL((k_mass_gemm<T, CT>), dim3(nb), dim3(NT), (mass_lds<T, CT>(S)), st, S, op, Mm, xt, yl,
                       stiff ? 1 : 0);
  timing_end(th, st);
}

template <>
void launch_apply_mfma<double>(const DevSpace& S, const OpCoef& op, const double* x, double* yl, hipStream_t st) {
  run<double, 2>(S, op, S.Dp64, S.Mm64, S.xt64, x, yl, st);
}
template <>
void launch_apply_mfma<float>(const DevSpace& S, const OpCoef& op, const float* x, float* yl, hipStream_t st) {
  run<float, 4>(S, op, S.Dp32, S.Mm32, S.xt32, x, yl, st);
}

}  // namespace tsem
