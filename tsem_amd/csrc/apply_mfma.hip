// Element operator on the matrix cores (gfx950 MFMA).
//
// For a tile of NE elements the stiffness action is two dense GEMMs through
// the derivative factors D_a (orthonormal coefficients of d_a phi_i):
//   U (3 Mp x NE) = D (3 Mp x K) X (K x NE)            X = gathered x_e
//   V_a = sum_b Ghat_ab(e) U_b                          per column (element)
//   Y (K x NE)   = D^T (K x 3 Mp) V (3 Mp x NE)
// with v_mfma_{f32,f64}_16x16x4.  Rows of each a-block are padded to Mp (a
// multiple of 16) so the three derivative components of one (mode, element)
// land in the same lane and the 3x3 mixing happens in registers.  X and V
// live in LDS; D is read from global memory (a few tens of KB, L1/L2
// resident) in the padded layout Dp[3 Mp][Cp], Cp = roundup(K, 16).
//
// MFMA 16x16x4 operand layout (one 64-lane wave): lane l supplies
// A[l % 16][l / 16] and B[l / 16][l % 16]; accumulator value j of lane l is
// C[4 (l / 16) + j][l % 16] for f32 and C[l / 16 + 4 j][l % 16] for f64
// (measured on gfx950).
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace tsem {

namespace {
constexpr int WAVES = 4;
constexpr int NT = 64 * WAVES;

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

// CT: number of 16-wide column tiles (NE = 16 CT elements per block)
template <typename T, int CT>
__global__ void __launch_bounds__(NT) k_apply_mfma(DevSpace S, OpCoef op, const T* __restrict__ Dp,
                                                   const T* __restrict__ x, T* __restrict__ yl) {
  using v4 = typename Mma<T>::v4;
  constexpr int NE = 16 * CT;
  extern __shared__ unsigned char smem_raw[];
  const int K = S.K;
  const int Kl = op.Kl > 0 ? op.Kl : K;
  const int Kdl = op.Kdl > 0 ? op.Kdl : S.Kd;
  const int Mp = S.Mp, Cp = S.Cp;                 // padded a-block rows, padded columns of Dp
  const int Kp = (Kl + 3) & ~3;                   // k extent of GEMM 1
  const int Mt = (Kdl + 15) >> 4;                 // m tiles per a-block in use
  const int Rt = (Kl + 15) >> 4;                  // output row tiles of GEMM 2
  T* X = reinterpret_cast<T*>(smem_raw);          // [Kp][NE]
  T* V = X + (size_t)((K + 3) & ~3) * NE;         // [3 Mp][NE]
  T* G = V + (size_t)3 * Mp * NE;                 // [NE][6]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int e0 = blockIdx.x * NE;
  const T cst = (T)op.stiff;

  // gather the tile: X[k][le]
  for (int s = tid; s < Kp * NE; s += NT) {
    int le = s % NE, k = s / NE, e = e0 + le;
    X[s] = (e < S.nt && k < Kl) ? gather_one<T>(S, e, k, x, op.constrained) : T(0);
  }
  for (int s = tid; s < NE * 6; s += NT) {
    int le = s / 6, e = e0 + le;
    G[s] = (e < S.nt) ? cst * (sizeof(T) == 8 ? (T)S.ghat[6 * e + s % 6] : (T)S.ghat32[6 * e + s % 6]) : T(0);
  }
  __syncthreads();

  const int lr = lane & 15, lk = lane >> 4;
  // GEMM 1 + mixing: wave w owns m tiles w, w + WAVES, ...
  for (int t = wave; t < Mt; t += WAVES) {
    v4 acc[3][CT];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int c = 0; c < CT; ++c) acc[a][c] = v4{0, 0, 0, 0};
    const T* d0 = Dp + (size_t)(0 * Mp + 16 * t + lr) * Cp + lk;
    const T* d1 = Dp + (size_t)(1 * Mp + 16 * t + lr) * Cp + lk;
    const T* d2 = Dp + (size_t)(2 * Mp + 16 * t + lr) * Cp + lk;
    for (int k = 0; k < Kp; k += 4) {
      T a0 = d0[k], a1 = d1[k], a2 = d2[k];
#pragma unroll
      for (int c = 0; c < CT; ++c) {
        T b = X[(k + lk) * NE + 16 * c + lr];
        acc[0][c] = Mma<T>::op(a0, b, acc[0][c]);
        acc[1][c] = Mma<T>::op(a1, b, acc[1][c]);
        acc[2][c] = Mma<T>::op(a2, b, acc[2][c]);
      }
    }
#pragma unroll
    for (int c = 0; c < CT; ++c) {
      const T* g = G + (16 * c + lr) * 6;
      T g0 = g[0], g1 = g[1], g2 = g[2], g3 = g[3], g4 = g[4], g5 = g[5];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        int m = 16 * t + Mma<T>::row(lk, j);
        T u0 = acc[0][c][j], u1 = acc[1][c][j], u2 = acc[2][c][j];
        T keep = (m < Kdl) ? T(1) : T(0);
        int col = 16 * c + lr;
        V[(size_t)(0 * Mp + m) * NE + col] = keep * (g0 * u0 + g3 * u1 + g4 * u2);
        V[(size_t)(1 * Mp + m) * NE + col] = keep * (g3 * u0 + g1 * u1 + g5 * u2);
        V[(size_t)(2 * Mp + m) * NE + col] = keep * (g4 * u0 + g5 * u1 + g2 * u2);
      }
    }
  }
  // m tiles beyond Mt are never written: clear them so GEMM 2 reads zeros
  for (int s = tid; s < 3 * (Mp - 16 * Mt) * NE; s += NT) {
    int per = (Mp - 16 * Mt) * NE;
    int a = s / per, r = s % per;
    V[(size_t)(a * Mp + 16 * Mt) * NE + r] = T(0);
  }
  __syncthreads();

  // GEMM 2: Y[i][e] = sum_r Dp[r][i] V[r][e]; wave w owns row tiles w, w + WAVES, ...
  T* Y = X;  // reuse of the (dead) X and V regions: [16 Rt][NE]
  v4 out[3][CT];
  int nown = 0;
  for (int t2 = wave; t2 < Rt; t2 += WAVES) {
    v4 acc[CT];
#pragma unroll
    for (int c = 0; c < CT; ++c) acc[c] = v4{0, 0, 0, 0};
    for (int a = 0; a < 3; ++a) {
      const T* dbase = Dp + (size_t)(a * Mp) * Cp + 16 * t2 + lr;
      const T* vbase = V + (size_t)(a * Mp) * NE;
      for (int k = 0; k < 16 * Mt; k += 4) {
        T av = dbase[(size_t)(k + lk) * Cp];
#pragma unroll
        for (int c = 0; c < CT; ++c) acc[c] = Mma<T>::op(av, vbase[(k + lk) * NE + 16 * c + lr], acc[c]);
      }
    }
    if (nown < 3) {
#pragma unroll
      for (int c = 0; c < CT; ++c) out[nown][c] = acc[c];
    }
    ++nown;
  }
  __syncthreads();  // all waves done reading X (gather) before it is reused for Y
  nown = 0;
  for (int t2 = wave; t2 < Rt; t2 += WAVES) {
#pragma unroll
    for (int c = 0; c < CT; ++c)
#pragma unroll
      for (int j = 0; j < 4; ++j) Y[(size_t)(16 * t2 + Mma<T>::row(lk, j)) * NE + 16 * c + lr] = out[nown][c][j];
    ++nown;
  }
  __syncthreads();
  // coalesced write yl[e][i] (zeros for modes beyond the level)
  for (int s = tid; s < NE * K; s += NT) {
    int le = s / K, i = s % K, e = e0 + le;
    if (e < S.nt) yl[(size_t)e * K + i] = (i < Kl) ? Y[(size_t)i * NE + le] : T(0);
  }
}

template <typename T, int CT>
static size_t lds_bytes(const DevSpace& S) {
  constexpr int NE = 16 * CT;
  return ((size_t)((S.K + 3) & ~3) * NE + (size_t)3 * S.Mp * NE + (size_t)NE * 6) * sizeof(T);
}

bool mfma_supported(const DevSpace& S) {
  // GEMM 2 keeps at most 3 row tiles per wave in registers: K <= 16*3*WAVES;
  // LDS must hold X, V and the tile's geometric factors
  return S.Dp64 != nullptr && S.K <= 16 * 3 * WAVES && lds_bytes<double, 2>(S) <= 160 * 1024 &&
         lds_bytes<float, 4>(S) <= 160 * 1024;
}

template <>
void launch_apply_mfma<double>(const DevSpace& S, const OpCoef& op, const double* x, double* yl, hipStream_t st) {
  constexpr int CT = 2;
  int nb = (S.nt + 16 * CT - 1) / (16 * CT);
  hipLaunchKernelGGL((k_apply_mfma<double, CT>), dim3(nb), dim3(NT), (lds_bytes<double, CT>(S)), st, S, op, S.Dp64, x, yl);
}
template <>
void launch_apply_mfma<float>(const DevSpace& S, const OpCoef& op, const float* x, float* yl, hipStream_t st) {
  constexpr int CT = 4;
  int nb = (S.nt + 16 * CT - 1) / (16 * CT);
  hipLaunchKernelGGL((k_apply_mfma<float, CT>), dim3(nb), dim3(NT), (lds_bytes<float, CT>(S)), st, S, op, S.Dp32, x, yl);
}

}  // namespace tsem
