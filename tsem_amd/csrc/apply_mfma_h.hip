// FP32 element operator on the fp16 matrix cores with two-term splitting.
//
// Every FP32 operand a is carried as a_hi + a_lo, both fp16, after scaling so
// that |a| <= 1 (per element for the element vectors, one power of two for
// the reference factors): a_hi = fp16(a), a_lo = fp16(a - a_hi).  A product
// a b is formed as a_hi b_hi + a_hi b_lo + a_lo b_hi on
// v_mfma_f32_16x16x32_f16 with FP32 accumulation; the dropped a_lo b_lo and
// the rounding of the parts leave a relative error of about 2^-22 per
// product (FP32 itself: 2^-24), at 16x the FP32 MFMA rate per instruction.
//
// k_elem_gemm_h: per block of NE = 16 CT elements, the entity DOFs of the
//   elements loaded from x into an FP32 LDS tile (consecutive threads on
//   consecutive DOFs of an entity), the element coefficients formed from it
//   (element descriptors, face table rows in LDS), scaled by the element's
//   max |x| and split into hi/lo fp16 in LDS in the MFMA B layout
//   [k / 8][element][8];
//   U = D X (GEMM 1), V = Ghat U per element (FP32, kept in registers),
//   exact per-element max |V_a| of each derivative component (LDS atomics),
//   Y = sum_a D_a^T V_a (GEMM 2) one component at a time: V_a / max split to
//   fp16 hi/lo into one of two LDS tiles (element-major, the B-operand
//   layout) and the accumulators rescaled between components; Y rescaled
//   and written element-major through an LDS transpose, or, from the same
//   LDS tile, as entity-slab entries (the element's contribution to each of
//   its global DOFs, written where the slab scatter reads it).
//
// MFMA 16x16x32 f16 operand layout (one 64-lane wave, measured on gfx950):
// lane l supplies A[l % 16][8 (l / 16) + t] and B[8 (l / 16) + t][l % 16],
// t = 0..7; accumulator value j of lane l is C[4 (l / 16) + j][l % 16].
#include <hip/hip_runtime.h>

#include <stdexcept>

#include "kernels.hpp"
#include "slab.hpp"

namespace tsem {

namespace {
constexpr int WAVES = 4;
constexpr int NT = 64 * WAVES;
constexpr int RMAX = 12;  // row tiles of GEMM 2 (K <= 192)

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ inline f4 mma(h8 a, h8 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0); }

// hi and lo must come from one FP32 rounding of the scaled value.  Left to
// itself the compiler rounds one of the two uses of fp16(v * inv) straight
// from the exact product (a mixed fp16 FMA) and the other from the FP32
// product; near fp16 midpoints those differ, and hi + lo is then one fp16
// ulp off.  The empty asm operand pins the FP32 product in a register.
__device__ inline float pin(float s) {
  asm volatile("" : "+v"(s));
  return s;
}

__device__ inline void split8(const float* v, float inv, h8& hi, h8& lo) {
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    float s = pin(v[t] * inv);
    _Float16 h = (_Float16)s;
    hi[t] = h;
    lo[t] = (_Float16)(s - (float)h);
  }
}

typedef _Float16 h4 __attribute__((ext_vector_type(4)));
__device__ inline h8 ld8(const uint16_t* p) { return *reinterpret_cast<const h8*>(p); }
__device__ inline void st8(uint16_t* p, h8 v) { *reinterpret_cast<h8*>(p) = v; }
__device__ inline void st4(uint16_t* p, h4 v) { *reinterpret_cast<h4*>(p) = v; }
}  // namespace

// 4 waves per SIMD: the phases are latency bound, occupancy pays more than
// registers.  The V tile holds one derivative component at a time (two
// tiles, 37 KB per block at p = 6), so four blocks fit a CU's LDS; on the
// 32^3, p = 6 mesh the kernel went 193 -> 185 us (p = 5: 151 -> 140 us) from
// three blocks with all three components in one tile, and at two blocks it
// took 234 us
template <int CT>
__global__ void __launch_bounds__(NT, 4) k_elem_gemm_h(DevSpace S, OpCoef op, const float* __restrict__ x,
                                                    float* __restrict__ yl, int slab) {
  constexpr int NE = 16 * CT;
  constexpr int VS = NE + 1;  // odd row stride of the FP32 LDS tiles: no bank conflicts
  extern __shared__ unsigned char smem_raw[];
  // one derivative component of V at a time: hi [NE][R2], lo [NE][R2] (phase C)
  uint16_t* Vh = reinterpret_cast<uint16_t*>(smem_raw);
  unsigned int* colmax = reinterpret_cast<unsigned int*>(smem_raw + S.hlds_v);  // [3][NE] float bits
  x += (size_t)blockIdx.y * op.ldx;  // column of a block apply
  yl += (size_t)blockIdx.y * op.ldl;
  const int K = S.K, K32 = S.K32, Mp = S.Mp, R2 = S.R2;
  const int Kl = op.Kl > 0 ? op.Kl : K;
  const int Kdl = op.Kdl > 0 ? op.Kdl : S.Kd;
  const int Mt = (Kdl + 15) >> 4;
  const int Rt = (Kl + 15) >> 4;
  const int K32l = (Kl + 31) & ~31;
  // the wave index is uniform over the wave: held in a scalar register, the
  // per-wave tables (the GEMM 2 row ends) are read with scalar loads, and no
  // vector-memory wait for them holds back the loads in flight
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lr = lane & 15, lk = lane >> 4;
  // each XCD takes a contiguous range of element blocks (xcd_block), so the
  // DOFs that neighbouring blocks share are gathered from, and their slab
  // lines written through, one L2
  const int e0 = xcd_block(blockIdx.x, gridDim.x) * NE;
  const int KB = K32 / 8;
  const float cst = (float)op.stiff * S.dinv;  // undo the factor scaling of GEMM 1
  for (int s = tid; s < 3 * NE; s += NT) colmax[s] = 0u;
  // slab form: the block's entity offsets and orientation words are loaded
  // now, into registers, and staged into LDS only after GEMM 2, so that
  // their latency passes behind the GEMMs
  constexpr int EOQ = (NE * 17 + NT - 1) / NT;
  int32_t eopre[EOQ];
  {
    const int nv0 = min(NE, S.nt - e0);
#pragma unroll
    for (int q = 0; q < EOQ; ++q) {
      const int s2 = tid + q * NT;
      eopre[q] = (slab && s2 < nv0 * 17) ? S.sl_eoff[(size_t)e0 * 17 + s2] : 0;
    }
  }
  // ---- phase 0: the block's element coefficients, gathered from x through
  // the element descriptors (phase 0a) and formed with the face table, scaled
  // by each element's max |x| and split to fp16 hi/lo into LDS in the MFMA B
  // layout [k / 8][element][8].  The region is the V tile's, dead until
  // phase B.
  const int R = 6 * S.ndg2, J = S.ftc_J;
  uint16_t* xhh = reinterpret_cast<uint16_t*>(smem_raw);
  uint16_t* xhl = xhh + (size_t)KB * NE * 8;
  ElemDesc* sd = reinterpret_cast<ElemDesc*>(xhl + (size_t)KB * NE * 8);
  int* sgm = reinterpret_cast<int*>(sd + NE);
  unsigned int* smx = reinterpret_cast<unsigned int*>(sgm + NE);  // per-element max |x|, float bits
  float* cw = reinterpret_cast<float*>(smx + NE);
  int* ck = reinterpret_cast<int*>(cw + (size_t)R * J);
  int* cn = ck + (size_t)R * J;
  int* sslot = cn + R;
  uint32_t* ssgn = reinterpret_cast<uint32_t*>(sslot + K);
  int* spos = reinterpret_cast<int*>(ssgn + 6);  // [K] entity order -> kind | geo << 2 | sub << 5
  const int n1 = S.M - 1, n2 = S.ndg2, FB = 4 + 6 * n1, IB = FB + 4 * n2;
  const int nvalid = min(NE, S.nt - e0);
  const int dg = S.sl_diag;
  {
    // every staging load is issued before the first of them is waited for
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
    float cwv[TQ];
    int ckv[TQ];
#pragma unroll
    for (int q = 0; q < TQ; ++q) {
      const int s2 = tid + NT * q;
      cwv[q] = s2 < R * J ? S.ftc_w32[s2] : 0.0f;
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
    if (tid < NE) { sgm[tid] = gmv; smx[tid] = 0u; }
#pragma unroll
    for (int q = 0; q < TQ; ++q) {
      const int s2 = tid + NT * q;
      if (s2 < R * J) { cw[s2] = cwv[q]; ck[s2] = ckv[q]; }
    }
    for (int s2 = tid + NT * TQ; s2 < R * J; s2 += NT) { cw[s2] = S.ftc_w32[s2]; ck[s2] = S.ftc_key[s2]; }
    if (tid < R) cn[tid] = cnv;
    for (int s2 = tid + NT; s2 < R; s2 += NT) cn[s2] = S.ftc_n[s2];
    if (tid < K) sslot[tid] = slv;
    for (int s2 = tid + NT; s2 < K; s2 += NT) sslot[s2] = S.slot[s2];
    if (tid < 6) ssgn[tid] = sgv;
    for (int p2 = tid; p2 < K; p2 += NT) {
      int kind = 3, geo = 0, sub = p2 - IB;
      if (p2 < 4) { kind = 0; geo = p2; sub = 0; }
      else if (p2 < FB) { kind = 1; geo = (p2 - 4) / n1; sub = p2 - 4 - geo * n1; }
      else if (p2 < IB) { kind = 2; geo = (p2 - FB) / n2; sub = p2 - FB - geo * n2; }
      // word of the entity's start in the element descriptor (v, e, f, then
      // t after fo), its mask bit, the DOF's place in the entity
      const int fw = kind == 0 ? geo : kind == 1 ? 4 + geo : kind == 2 ? 10 + geo : 18;
      spos[p2] = fw | (kind == 3 ? 14 : fw) << 5 | sub << 9;
    }
  }
  __syncthreads();
  // ---- phase 0a: the block's entity DOFs, masked, into an FP32 LDS tile in
  // entity order (4 vertices, 6 edges of n1, 4 faces of n2, the interior),
  // [pos][element] with the element index swizzled by the row so that the
  // writes here and the reads below are free of bank conflicts.  One DOF per
  // thread and step, consecutive threads on consecutive DOFs of an element:
  // the DOFs of an entity are contiguous in x, so a wave's loads touch a few
  // lines per entity.  Loaded one element per lane, every lane touched a line
  // of its own: the kernel took 184 us at p = 6 on the 32^3 mesh against
  // 173 us with the tile, 350 at p = 7 (320) and 140 at p = 5 (141).  The
  // tile lies in the split tile's region, dead until the split below.
  float* rawt = reinterpret_cast<float*>(smem_raw);
  {
    constexpr int DW = sizeof(ElemDesc) / 4;
    static_assert(offsetof(ElemDesc, t) == 72, "descriptor words: v 0-3, e 4-9, f 10-13, fo 14-17, t 18");
    constexpr int RQ = NE * (CT == 4 ? 96 : 192) / NT;
    const int nk = nvalid * K;
    const int dle = NT / K, dpos = NT - dle * K;
    int le = tid / K, pos = tid - le * K;
    int ad[RQ], ri[RQ];
#pragma unroll
    for (int q = 0; q < RQ; ++q) {
      ad[q] = -1;
      ri[q] = -1;
      if (tid + NT * q < nk) {
        // no branch on the entity kind, which differs across the lanes: the
        // start is one descriptor word (the edge flag in bit 31 cleared).
        // With a branch per kind, p = 6 on the 32^3 mesh took 169 us against
        // 157 us without, p = 5 137 (130), p = 7 321 (306)
        const int code = spos[pos];
        const int a2 = (reinterpret_cast<const int32_t*>(sd)[le * DW + (code & 31)] & 0x7fffffff) + (code >> 9);
        ad[q] = ((sgm[le] >> ((code >> 5) & 15)) & 1) ? -1 : a2;
        ri[q] = pos * NE + (le ^ (pos & (NE - 1)));
      }
      pos += dpos;
      le += dle;
      if (pos >= K) { pos -= K; ++le; }
    }
    // steps past the block's last DOF are skipped whole (a uniform branch)
    float rv[RQ];
#pragma unroll
    for (int q = 0; q < RQ; ++q) rv[q] = NT * q < nk ? x[ad[q] < 0 ? 0 : ad[q]] : 0.0f;
#pragma unroll
    for (int q = 0; q < RQ; ++q)
      if (NT * q < nk && ri[q] >= 0) rawt[ri[q]] = ad[q] < 0 ? 0.0f : rv[q];
  }
  __syncthreads();
  // three groups of 8 modes per thread (NE * K32 / 8 = 768 groups: K32 <= 96
  // with CT = 4, <= 192 with CT = 2), element fastest across the threads
  constexpr int GQ = 3;
  static_assert(NE * (CT == 4 ? 12 : 24) <= GQ * NT, "three groups of 8 modes per thread");
  const int KBl = K32l / 8;
  float xv[GQ][8];
  auto rawv = [&](int le, int pos) -> float { return rawt[pos * NE + (le ^ (pos & (NE - 1)))]; };
  auto coef = [&](int le, int k) -> float {
    const int code = CT == 4 ? S.slot[k] : sslot[k];
    const int kind = code & 3, geo = (code >> 2) & 7, sub = code >> 5;
    if (kind == 0) return rawv(le, geo);
    if (kind == 1) {
      const float v = rawv(le, 4 + geo * n1 + sub);
      return (sd[le].e[geo] < 0 && (sub & 1)) ? -v : v;
    }
    if (kind == 3) return rawv(le, IB + sub);
    const int fo = sd[le].fo[geo];
    if ((dg >> (fo & 7)) & 1) {
      const float v = rawv(le, FB + geo * n2 + sub);
      return (sub < 32 && ((ssgn[fo & 7] >> sub) & 1)) ? -v : v;
    }
    const int row = (fo & 7) * S.ndg2 + sub;
    float acc = 0.0f;
    const int nterm = cn[row];
    for (int j = 0; j < nterm; ++j) {
      const int key = ck[row * J + j];
      const int ie = key & 3, idx = key >> 2;
      float w = cw[row * J + j];
      float xq;
      if (ie < 3) {
        const int le2 = (fo >> (3 + 3 * ie)) & 7;
        if (((fo >> (12 + ie)) & 1) && (idx & 1)) w = -w;
        xq = rawv(le, 4 + le2 * n1 + idx);
      } else {
        xq = rawv(le, FB + geo * n2 + idx);
      }
      acc += xq * w;
    }
    return acc;
  };
#pragma unroll
  for (int g = 0; g < GQ; ++g) {
    // with NE = 64 a wave holds one group of 8 modes over its 64 elements:
    // the mode, its slot code (a scalar load) and its entity kind are uniform
    // over the wave, and so are the branches on them (p = 6 on the 32^3
    // mesh: 174 -> 169 us against per-lane codes from LDS; p = 5: 141 -> 137)
    const int gi = tid + NT * g, kb = CT == 4 ? wave + (NT / NE) * g : gi / NE, le = CT == 4 ? lane : gi - kb * NE;
    float m = 0.0f;
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int k = 8 * kb + t;
      xv[g][t] = (kb < KBl && le < nvalid && k < Kl) ? coef(le, k) : 0.0f;
      m = fmaxf(m, fabsf(xv[g][t]));
    }
    if (kb < KBl) atomicMax(&smx[le], __float_as_uint(m));
  }
  __syncthreads();
#pragma unroll
  for (int g = 0; g < GQ; ++g) {
    const int gi = tid + NT * g, kb = CT == 4 ? wave + (NT / NE) * g : gi / NE, le = CT == 4 ? lane : gi - kb * NE;
    if (kb < KBl) {
      const float m = __uint_as_float(smx[le]);
      h8 hi, lo;
      split8(xv[g], m > 0.0f ? 1.0f / m : 0.0f, hi, lo);
      st8(xhh + ((size_t)kb * NE + le) * 8, hi);
      st8(xhl + ((size_t)kb * NE + le) * 8, lo);
    }
  }
  __syncthreads();

  // ---- phase A: GEMM 1 + mixing; wave w owns m tiles w + WAVES tp, tp <
  // TPW; V stays in registers (vv) until the per-element maxima are complete
  constexpr int TPW = 4 / CT;  // m tiles per wave: Mt <= 4 with CT = 4, <= 8 with CT = 2
  float vv[TPW][CT][3][4];
#pragma unroll
  for (int tp = 0; tp < TPW; ++tp) {
    const int t = wave + WAVES * tp;
    if (t >= Mt) break;
    f4 acc[3][CT];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int c = 0; c < CT; ++c) acc[a][c] = f4{0, 0, 0, 0};
    const int kb1 = S.kbeg1h[t];
    for (int kk = kb1; kk < K32l; kk += 32) {
      h8 ah[3], al[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        size_t o = ((size_t)(a * Mp + 16 * t + lr) * K32 + kk + 8 * lk);
        ah[a] = ld8(S.Dh1hi + o);
        al[a] = ld8(S.Dh1lo + o);
      }
      const size_t ko = (size_t)(kk / 8 + lk) * NE * 8;
#pragma unroll
      for (int c = 0; c < CT; ++c) {
        const size_t o = ko + (size_t)(16 * c + lr) * 8;
        h8 bh = ld8(xhh + o), bl = ld8(xhl + o);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          acc[a][c] = mma(ah[a], bh, acc[a][c]);
          acc[a][c] = mma(ah[a], bl, acc[a][c]);
          acc[a][c] = mma(al[a], bh, acc[a][c]);
        }
      }
    }
#pragma unroll
    for (int c = 0; c < CT; ++c) {
      int col = 16 * c + lr, e = e0 + col;
      float g0 = 0, g1 = 0, g2 = 0, g3 = 0, g4 = 0, g5 = 0;
      if (e < S.nt) {
        const float* g = S.ghat32 + 6 * (size_t)e;
        float f = cst * __uint_as_float(smx[col]);
        g0 = f * g[0]; g1 = f * g[1]; g2 = f * g[2]; g3 = f * g[3]; g4 = f * g[4]; g5 = f * g[5];
      }
      float vmax0 = 0.0f, vmax1 = 0.0f, vmax2 = 0.0f;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        int m = 16 * t + 4 * lk + j;
        float u0 = acc[0][c][j], u1 = acc[1][c][j], u2 = acc[2][c][j];
        float keep = (m < Kdl) ? 1.0f : 0.0f;
        float v0 = keep * (g0 * u0 + g3 * u1 + g4 * u2);
        float v1 = keep * (g3 * u0 + g1 * u1 + g5 * u2);
        float v2 = keep * (g4 * u0 + g5 * u1 + g2 * u2);
        vv[tp][c][0][j] = v0;
        vv[tp][c][1][j] = v1;
        vv[tp][c][2][j] = v2;
        vmax0 = fmaxf(vmax0, fabsf(v0));
        vmax1 = fmaxf(vmax1, fabsf(v1));
        vmax2 = fmaxf(vmax2, fabsf(v2));
      }
      atomicMax(&colmax[col], __float_as_uint(vmax0));
      atomicMax(&colmax[NE + col], __float_as_uint(vmax1));
      atomicMax(&colmax[2 * NE + col], __float_as_uint(vmax2));
    }
  }
  __syncthreads();

  // ---- phase B: the tile rows from 16 Mt to Mh are zero for GEMM 2; no
  // component writes them, so they are cleared once
  const int Mh = S.Mh;
  {
    const int zr = (Mh - 16 * Mt) / 4;  // groups of 4 zero rows
    for (int s = tid; s < 4 * NE * zr; s += NT) {
      const int g = s % zr, row = s / zr;  // rows of both buffers: hi, lo, hi, lo
      st4(Vh + (size_t)row * R2 + 16 * Mt + 4 * g, h4{0, 0, 0, 0});
    }
  }

  // ---- phase C: GEMM 2, Y[i][e] = sum_a sum_r D_a[r][i] V_a[r][e], one
  // derivative component a at a time from its LDS tile.  Each wave owns
  // RJ row tiles and NC = 2 column tiles: with CT = 2 every column and the
  // row tiles w + WAVES j; with CT = 4 (K <= 96, at most 6 row tiles) the
  // column tiles 2 (w & 1) + {0, 1} and the row tiles (w >> 1) + 2 j, so the
  // six row tiles of p = 6 give every wave the same MFMA count (rows split
  // 2-2-1-1 over whole column sets gave two waves twice the work).  Per
  // 32-row step a wave issues all its factor loads (global) and V loads
  // (LDS) at once, then the MFMAs
  constexpr int RJ = (RMAX + WAVES - 1) / WAVES;
  constexpr int NC = 2;
  static_assert(CT == 2 || CT == 4, "two or four column tiles");
  auto tile2 = [&](int j) { return CT == 4 ? (wave >> 1) + 2 * j : wave + WAVES * j; };
  auto ctile = [&](int cc) { return CT == 4 ? 2 * (wave & 1) + cc : cc; };
  f4 acc[RJ][NC];
#pragma unroll
  for (int j = 0; j < RJ; ++j)
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[j][c] = f4{0, 0, 0, 0};
  int kend[RJ];
  int kmax = 0;
#pragma unroll
  for (int j = 0; j < RJ; ++j) {
    int t2 = tile2(j);
    kend[j] = (t2 < Rt) ? min(S.kend2h[t2], 16 * Mt) : 0;
    kmax = max(kmax, kend[j]);
  }
  // acc carries Y / run[c] for the lane's element column: each component is
  // split with its own per-element maximum, and acc is brought to the new
  // scale before the component is added (run stays 0 until a component with
  // a nonzero maximum)
  float run[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) run[c] = 0.0f;
  // component a of V / its maxima, split to fp16 hi/lo from the registers,
  // into tile buffer a & 1 (components 0 and 1 first, 2 once 0 is read)
  auto put = [&](int a) {
    uint16_t* T = Vh + (size_t)(a & 1) * 2 * NE * R2;
#pragma unroll
    for (int tp = 0; tp < TPW; ++tp) {
      const int t = wave + WAVES * tp;
      if (t >= Mt) break;
#pragma unroll
      for (int c = 0; c < CT; ++c) {
        const int col = 16 * c + lr;
        const float m = __uint_as_float(colmax[a * NE + col]);
        const float inv = m > 0.0f ? 1.0f / m : 0.0f;
        h4 hi, lo;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const float sv = pin(vv[tp][c][a][j] * inv);  // one FP32 rounding for hi and lo
          const _Float16 h = (_Float16)sv;
          hi[j] = h;
          lo[j] = (_Float16)(sv - (float)h);
        }
        const size_t o = (size_t)16 * t + 4 * lk;
        st4(T + (size_t)col * R2 + o, hi);
        st4(T + (size_t)(NE + col) * R2 + o, lo);
      }
    }
  };
  auto gemm = [&](int a) {
    const uint16_t* T = Vh + (size_t)(a & 1) * 2 * NE * R2;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const float m = __uint_as_float(colmax[a * NE + 16 * ctile(c) + lr]);
      if (m > 0.0f) {
        const float f = run[c] / m;
#pragma unroll
        for (int j = 0; j < RJ; ++j) acc[j][c] *= f;
        run[c] = m;
      }
    }
    for (int mm = 0; mm < kmax; mm += 32) {
      const int rv = mm + 8 * lk, rr = a * Mh + rv;
      h8 ah[RJ], al[RJ], bh[NC], bl[NC];
#pragma unroll
      for (int j = 0; j < RJ; ++j) {
        int t2 = tile2(j);
        int i = 16 * (t2 < Rt ? t2 : 0) + lr;
        size_t o = (size_t)i * S.R2g + rr;
        ah[j] = ld8(S.Dh2hi + o);
        al[j] = ld8(S.Dh2lo + o);
      }
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        int col = 16 * ctile(c) + lr;
        bh[c] = ld8(T + (size_t)col * R2 + rv);
        bl[c] = ld8(T + (size_t)(NE + col) * R2 + rv);
      }
#pragma unroll
      for (int j = 0; j < RJ; ++j) {
        if (mm < kend[j]) {
#pragma unroll
          for (int c = 0; c < NC; ++c) {
            acc[j][c] = mma(ah[j], bh[c], acc[j][c]);
            acc[j][c] = mma(ah[j], bl[c], acc[j][c]);
            acc[j][c] = mma(al[j], bh[c], acc[j][c]);
          }
        }
      }
    }
  };
  put(0);
  put(1);
  __syncthreads();
  gemm(0);
  __syncthreads();  // buffer 0 read by every wave
  put(2);
  gemm(1);
  __syncthreads();  // component 2 in place, buffer 1 read by every wave
  gemm(2);
  __syncthreads();  // every wave is done reading Vh

  // ---- phase D: rescale and write Y element-major through an LDS transpose
  float* Y = reinterpret_cast<float*>(smem_raw);  // [K][VS]
  const int nval = min(NE, S.nt - e0);
  // slab form: Y rows in slab order (sl_perm); the block's entity offsets
  // and the transposed face rows are staged behind Y, in the dead fp16 V
  // tile below the maxima (slab_h_fits), while Y is written, so that the
  // epilogue reads LDS only
  const int JT = S.sl_JT, nrow = S.sl_nrow;
  int32_t* seo = nullptr;   // [NE][17] entity offsets, orientation words
  float* stw = nullptr;     // [nrow][JT] weights
  uint16_t* stk = nullptr;  // [nrow][JT] Y offsets of the terms' face modes (d VS)
  uint16_t* stn = nullptr;  // [nrow] term counts
  if (slab) {
    const int nrj = nrow * JT;
    seo = reinterpret_cast<int32_t*>(smem_raw + slab_tab_offset(K, NE));
    stw = reinterpret_cast<float*>(seo + NE * 17);
    stk = reinterpret_cast<uint16_t*>(stw + nrj);
    stn = stk + nrj;
#pragma unroll
    for (int q = 0; q < EOQ; ++q)
      if (tid + q * NT < NE * 17) seo[tid + q * NT] = eopre[q];
    for (int s = tid; s < nrj; s += NT) { stw[s] = S.sl_tw32[s]; stk[s] = (uint16_t)((S.sl_tk[s] & 255) * VS); }
    for (int s = tid; s < nrow; s += NT) stn[s] = (uint16_t)(S.sl_tk[s * JT] >> 8);
    int32_t* work = reinterpret_cast<int32_t*>(smem_raw + slab_tab_offset(K, NE) + slab_tab_bytes(S, NE));
    if (tid < 2) work[tid] = 0;  // work-list counts
    if (tid < 6) work[2 + tid] = (int32_t)S.sl_sgn[tid];
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const int col = 16 * ctile(c) + lr;
    const float m = run[c] * S.dinv;
#pragma unroll
    for (int j = 0; j < RJ; ++j) {
      int t2 = tile2(j);
      if (t2 < Rt) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          int i = 16 * t2 + 4 * lk + q;
          if (i < K) Y[(size_t)(slab ? S.sl_perm[i] : i) * VS + col] = acc[j][c][q] * m;
        }
      }
    }
  }
  __syncthreads();
  if (slab) {  // full order only: every mode of Y is set
    // The epilogue is bound by vector instructions (a wave64 instruction
    // takes 4 cycles of a 16-lane SIMD), so it runs in two passes.  Pass A:
    // every entry that is a copy -- vertices, interiors, DOFs of faces in a
    // diagonal variant (+- the element's own face mode; the identity of
    // variant 0 among them) and edges whose two faces are diagonal (own mode,
    // edge sign) -- with one entry per thread, decoded once, walking the
    // elements; consecutive threads hold consecutive DOFs of one entity, so
    // stores come in runs.  Meanwhile the faces in other variants and the
    // edges next to them go to LDS work lists (their order does not matter:
    // every item writes its own entries); pass B sums their transposed rows
    // term by term in table order (as slab_apply_host does on the host).
    const int n1 = S.sl_n1, n2 = S.sl_n2, n3 = K - 4 - 6 * n1 - 4 * n2;
    const int FB = 4 + 6 * n1;  // first face row of Y
    const int dg = S.sl_diag;
    int32_t* work = reinterpret_cast<int32_t*>(smem_raw + slab_tab_offset(K, NE) + slab_tab_bytes(S, NE));
    const int32_t* sgn = work + 2;
    int32_t* itemF = work + 8;        // [NE * 4] le * 4 + f
    int32_t* itemE = itemF + NE * 4;  // [NE * 6] le * 6 + e
    auto faces_of = [](int e, int& fa, int& fb) {  // the two faces holding edge e, lower first
      const int a = e < 3 ? 0 : e < 5 ? 1 : 2, b = e < 3 ? e + 1 : e < 5 ? e - 1 : 3;
      const int rest = 15 & ~((1 << a) | (1 << b));
      fa = rest & 1 ? 0 : rest & 2 ? 1 : 2;
      fb = rest & 8 ? 3 : rest & 4 ? 2 : 1;
    };
    if (tid < 4 * nval) {
      yl[seo[(tid >> 2) * 17 + (tid & 3)]] = Y[(tid & 3) * VS + (tid >> 2)];
      const int var = (seo[(tid >> 2) * 17 + 15] >> (6 * (tid & 3))) & 7;
      if (!((dg >> var) & 1)) itemF[atomicAdd(&work[0], 1)] = tid;
    }
    for (int s = tid; s < 6 * nval; s += NT) {
      const int le = s / 6, e = s - 6 * le;
      int fa, fb;
      faces_of(e, fa, fb);
      const int32_t ow = seo[le * 17 + 15];
      if (!((dg >> ((ow >> (6 * fa)) & 7)) & (dg >> ((ow >> (6 * fb)) & 7)) & 1)) itemE[atomicAdd(&work[1], 1)] = s;
    }
    if (n1 > 0) {
      const int EE = 6 * n1, G = NT / EE, g = tid / EE, r = tid - g * EE;
      if (g < G) {
        const int e = r / n1, l = r - e * n1;
        int fa, fb;
        faces_of(e, fa, fb);
        const int yo = (4 + e * n1 + l) * VS;
        for (int le = g; le < nval; le += G) {
          const int32_t* eo = seo + le * 17;
          const int32_t ow = eo[15];
          if (!((dg >> ((ow >> (6 * fa)) & 7)) & (dg >> ((ow >> (6 * fb)) & 7)) & 1)) continue;
          const float v = Y[yo + le];
          yl[eo[4 + e] + l] = (((ow >> (24 + e)) & 1) && (l & 1)) ? -v : v;
        }
      }
    }
    if (n2 > 0) {
      const int EF = 4 * n2, G = NT / EF, g = tid / EF, r = tid - g * EF;
      if (g < G) {
        const int f = r / n2, i = r - f * n2;
        const int yo = (FB + f * n2 + i) * VS;
        for (int le = g; le < nval; le += G) {
          const int32_t* eo = seo + le * 17;
          const int var = (eo[15] >> (6 * f)) & 7;
          if (!((dg >> var) & 1)) continue;
          const float v = Y[yo + le];
          yl[eo[10 + f] + i] = (i < 32 && ((sgn[var] >> i) & 1)) ? -v : v;
        }
      }
    }
    if (n3 > 0) {
      const int G = NT / n3, g = tid / n3, i = tid - g * n3;
      if (g < G)
        for (int le = g; le < nval; le += G) yl[seo[le * 17 + 14] + i] = Y[(K - n3 + i) * VS + le];
    }
    __syncthreads();
    // pass B: faces in a non-diagonal variant, then the edges next to them
    const int nF = work[0], nE = work[1];
    for (int s = tid; s < nF * n2; s += NT) {
      const int it = s / n2, i = s - it * n2;
      const int le = itemF[it] >> 2, f = itemF[it] & 3;
      const int32_t* eo = seo + le * 17;
      const int row = ((eo[15] >> (6 * f)) & 7) * n2 + i;
      const float* w = stw + row * JT;
      const uint16_t* k = stk + row * JT;
      const float* Yf = Y + (FB + f * n2) * VS + le;
      float v = 0.0f;
      for (int t = 0; t < stn[row]; ++t) v += w[t] * Yf[k[t]];
      yl[eo[10 + f] + i] = v;
    }
    for (int s = tid; s < nE * n1; s += NT) {
      const int it = s / n1, l = s - it * n1;
      const int le = itemE[it] / 6, e = itemE[it] - 6 * le;
      const int32_t* eo = seo + le * 17;
      const int32_t ow = eo[15], ow2 = eo[16];
      int fh[2];
      faces_of(e, fh[0], fh[1]);
      float v = Y[(4 + e * n1 + l) * VS + le];
      if (((ow >> (24 + e)) & 1) && (l & 1)) v = -v;
      float part[2] = {0.0f, 0.0f};
      for (int h = 0; h < 2; ++h) {
        const int f = fh[h], var = (ow >> (6 * f)) & 7;
        if ((dg >> var) & 1) continue;
        const int q = (ow2 >> (4 * e + 2 * h)) & 3;
        const int row = 6 * n2 + (var * 3 + q) * n1 + l;
        const float* w = stw + row * JT;
        const uint16_t* k = stk + row * JT;
        const float* Yf = Y + (FB + f * n2) * VS + le;
        float p = 0.0f;
        for (int t = 0; t < stn[row]; ++t) p += w[t] * Yf[k[t]];
        part[h] = (((ow >> (6 * f + 3 + q)) & 1) && (l & 1)) ? -p : p;
      }
      yl[eo[4 + e] + l] = (v + part[0]) + part[1];
    }
    return;
  }
  for (int s = tid; s < nval * K; s += NT) {
    int le = s / K, i = s % K;
    yl[(size_t)(e0 + le) * K + i] = (i < Kl && i < 16 * Rt) ? Y[(size_t)i * VS + le] : 0.0f;
  }
}

int gemm_h_ct(const DevSpace& S) { return S.Mp <= 64 ? 4 : 2; }

// end of the slab data staged behind the Y tile (slab_tab_offset): the
// tables (slab_tab_bytes), then the epilogue's work-list counts, sign masks
// and work lists (8 + NE * 10 ints)
static size_t slab_tab_end(const DevSpace& S) {
  const int NE = 16 * gemm_h_ct(S);
  return slab_tab_offset(S.K, NE) + slab_tab_bytes(S, NE) + (size_t)(8 + NE * 10) * 4;
}

// the tables fit below the maxima, and each entity kind's DOFs of one
// element fit in one block of threads (the epilogue gives each thread one)
bool slab_h_fits(const DevSpace& S) {
  const int n3 = S.K - 4 - 6 * S.sl_n1 - 4 * S.sl_n2;
  const int VS = 16 * gemm_h_ct(S) + 1;
  return slab_tab_end(S) <= S.hlds_v && 6 * S.sl_n1 <= NT && 4 * S.sl_n2 <= NT && n3 <= NT &&
         (S.sl_n2 - 1) * VS < 65536;
}

size_t gemm_h_lds(const DevSpace& S) { return gemm_h_colmax_offset(S) + (size_t)3 * 16 * gemm_h_ct(S) * 4; }

// bytes of the gather phase's LDS (split tile, which the entity tile of
// phase 0a shares, descriptors, masks and maxima, face table rows, slot
// codes, sign masks, entity codes), below the maxima of V
static size_t gather_lds_end(const DevSpace& S) {
  const size_t NE = 16 * gemm_h_ct(S), R = 6 * (size_t)S.ndg2, J = S.ftc_J;
  return 4 * (size_t)S.K32 * NE + NE * sizeof(ElemDesc) + 8 * NE + R * J * 8 + R * 4 + (size_t)S.K * 8 + 24;
}

// LDS byte offset of the per-element maxima: after the largest of the
// gather phase's data, one component's fp16 V tile, the Y tile and the slab
// data behind it
size_t gemm_h_colmax_offset(const DevSpace& S) {
  const size_t NE = 16 * gemm_h_ct(S);
  const size_t vh = (size_t)4 * NE * S.R2 * 2;  // two tiles, hi and lo
  const size_t y = (size_t)S.K * (NE + 1) * 4;
  size_t m = vh > y ? vh : y;
  m = m > gather_lds_end(S) ? m : gather_lds_end(S);
  if (S.sl_meta) m = m > slab_tab_end(S) ? m : slab_tab_end(S);
  return (m + 15) & ~(size_t)15;
}

bool mfma_h_supported(const DevSpace& S) {
  return S.Dh1hi != nullptr && S.K <= 16 * RMAX && S.Mp <= 128 && S.K32 <= 192 &&
         S.K32 <= (gemm_h_ct(S) == 4 ? 96 : 192) && S.Mp / 16 <= WAVES * (4 / gemm_h_ct(S)) &&
         gemm_h_lds(S) <= 160 * 1024 && gather_lds_end(S) <= S.hlds_v;
}

void launch_apply_mfma_h(const DevSpace& S, const OpCoef& op, const float* x, float* yl, bool slab, hipStream_t st) {
  if (slab && (op.Kl != 0 || !S.sl_meta)) throw std::logic_error("slab results need the full order and slab tables");
  const bool timed = op.timed();
  const int CT = gemm_h_ct(S);
  const int NE = 16 * CT;
  int nb = (S.nt + NE - 1) / NE;
  // one kernel: gather, both GEMMs, mixing, slab epilogue (timed as the element kernel)
  int th = timed ? timing_begin(TK_ELEM32, st) : -1;
  if (op.ncol < 1) throw std::logic_error("block apply needs at least one column");
  const dim3 grid(nb, op.ncol);
  if (CT == 4)
    hipLaunchKernelGGL(k_elem_gemm_h<4>, grid, dim3(NT), gemm_h_lds(S), st, S, op, x, yl, (int)slab);
  else
    hipLaunchKernelGGL(k_elem_gemm_h<2>, grid, dim3(NT), gemm_h_lds(S), st, S, op, x, yl, (int)slab);
  timing_end(th, st);
}

}  // namespace tsem
