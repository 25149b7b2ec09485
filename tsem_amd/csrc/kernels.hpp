// Device-side data of a space and the launchers of the HIP kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "space.hpp"

namespace tsem {

// Gathered element vectors feeding the GEMM live in a tiled layout: tiles of
// TILE elements, mode-major inside a tile: xt[tile][k][TILE].  Element
// results are element-major: yl[e][i].
constexpr int TILE = 32;

// The blocks of one launch are dispatched round-robin over the 8 XCDs (block
// b runs on XCD b % 8), and each XCD has its own L2.  xcd_block maps the
// hardware block index to a logical one so that every XCD walks one
// contiguous range of logical blocks: indirect accesses made by neighbouring
// work then hit lines already in that XCD's L2.  A bijection on [0, nb).
constexpr int NXCD = 8;
__device__ inline int xcd_block(int b, int nb) {
  int q = nb / NXCD, r = nb % NXCD, x = b % NXCD;
  return x * q + (x < r ? x : r) + b / NXCD;
}
inline size_t tiled_index(int e, int k, int rows) {
  return (size_t)(e / TILE) * rows * TILE + (size_t)k * TILE + (e % TILE);
}
// tiles allocated: whole blocks of 64 elements (the widest GEMM block)
inline int tile_count(int nt) { return 2 * ((nt + 63) / 64); }

// Everything the kernels read about a space (device pointers).
struct DevSpace {
  int nt = 0, K = 0, Kd = 0, M = 0, nface = 0;
  int ndof = 0;
  int sc_chunk_max = 0;              // most scatter terms of any 256-row chunk
  const ElemDesc* desc = nullptr;
  const int32_t* slot = nullptr;     // per local mode: kind | geo << 2 | sub << 5
  const int32_t* ft_ptr = nullptr;   // face table CSR over (variant, face dof)
  const int32_t* ft_key = nullptr;   // ie | idx << 2
  const double* ft_w = nullptr;
  // the rows of the face table this space uses (variant v < 6, face mode
  // d < ndg2), padded to ftc_J terms: row v ndg2 + d holds ftc_n terms
  // ftc_key (ie | idx << 2) and ftc_w / ftc_w32; small enough for LDS
  int ndg2 = 0, ftc_J = 0;
  const int32_t* ftc_n = nullptr;
  const int32_t* ftc_key = nullptr;
  const double* ftc_w = nullptr;
  const float* ftc_w32 = nullptr;
  const double* ghat = nullptr;      // nt x 6
  const float* ghat32 = nullptr;
  const double* sixv = nullptr;      // nt
  const float* sixv32 = nullptr;
  const double* dT = nullptr;        // K x 3Kd  (derivative factors, transposed)
  const double* dN = nullptr;        // 3Kd x K
  const float* dT32 = nullptr;
  const float* dN32 = nullptr;
  const double* mT = nullptr;        // K x K (mass factor transposed)
  const double* mN = nullptr;
  const float* mT32 = nullptr;
  const float* mN32 = nullptr;
  const int32_t* sc_ptr = nullptr;   // ndof + 1
  const int32_t* sc_code = nullptr;  // slot | bit30 negative | bit31 weighted
  const double* sc_w = nullptr;
  const float* sc_w32 = nullptr;
  const double* diag = nullptr;      // assembled operator diagonal
  const double* invdiag = nullptr;
  const float* invdiag32 = nullptr;
  const uint8_t* cmask = nullptr;    // 1 where the DOF is eliminated (Dirichlet)
  const int32_t* gmask = nullptr;
  // 3-D quadrature of the space (load vector, point values, M_V): nq points,
  // weights wq (sum 1 over the reference tet), phi[q][i], phiT[i][q]; vpot:
  // potential at the points, [nt][nq] (M_V, set by the caller); xe: vertex
  // coordinates [nt][4][3]; pts: reference coordinates of the points [nq][3]
  int nq = 0;
  const double* wq = nullptr;
  const double* phi = nullptr;
  const double* phiT = nullptr;
  const double* dphi = nullptr;   // reference gradients at the points: [nq][K][3]
  const double* vpot = nullptr;
  const double* xe = nullptr;
  const double* pts = nullptr;    // per element: eliminated geometries (v0..3, e0..5, f0..3, interior)
  // padded derivative factors for the MFMA path: Dp[3 Mp][Cp]
  int Mp = 0, Cp = 0, Kp = 0, ntile = 0;
  const double* Dp64 = nullptr;
  const float* Dp32 = nullptr;
  // reference mass matrix mfac^T mfac, Cp x Cp (zero padded), for the MFMA
  // mass kernel
  const double* Mm64 = nullptr;
  const float* Mm32 = nullptr;
  // per 16-row tile: first column of GEMM 1, last row (exclusive) of GEMM 2
  int kbeg1[16] = {0};
  int kend2[16] = {0};
  // fp16-split FP32 operator: factors scaled by 1/dinv and split hi/lo,
  // Dh1[3][Mp][K32] (GEMM 1), Dh2[K16][3 Mh] (GEMM 2, D transposed, each
  // component padded to Mh = roundup(Kd, 32) rows, row stride R2g = 3 Mh);
  // the element kernel's LDS tile of one component of V has row stride
  // R2 = Mh + 8 (the element vectors are gathered and split inside it)
  int K32 = 0, Mh = 0, R2 = 0, R2g = 0;
  size_t hlds_v = 0;  // LDS byte offset of the per-element maxima
  float dinv = 1.0f;
  int kbeg1h[16] = {0};
  int kend2h[16] = {0};
  const uint16_t *Dh1hi = nullptr, *Dh1lo = nullptr, *Dh2hi = nullptr, *Dh2lo = nullptr;
  // tiled scratch for gathered element vectors
  double* xt64 = nullptr;
  float* xt32 = nullptr;
  // entity-slab layout of element results (slab.hpp); present when sl_meta
  // is set: per DOF meta (offset in its chunk | count << 20 | dim << 30),
  // per chunk of 256 DOFs its span (start, length), per element 15 entity
  // offsets and two orientation words, local mode -> slab row, transposed
  // face rows (slab.hpp)
  int sl_n1 = 0, sl_n2 = 0, sl_JT = 1, sl_nrow = 1, sl_chunk_max = 0;
  int sl_diag = 0;                    // diagonal face variants (bit v; 0 without slab tables)
  uint32_t sl_sgn[6] = {0, 0, 0, 0, 0, 0};  // their minus signs per face DOF
  const int32_t* sl_meta = nullptr;
  const int32_t* sl_chunk = nullptr;
  const int32_t* sl_eoff = nullptr;  // [nt][17]
  const int32_t* sl_perm = nullptr;  // [K] slab row of local mode i
  const uint16_t* sl_tk = nullptr;
  const double* sl_tw = nullptr;
  const float* sl_tw32 = nullptr;
};

// A Chebyshev step fused into the scatter of an apply (p-multigrid
// smoothing): in place of y = A x the scatter forms d = c1 d + c2 Dinv (b - A x)
// and x += d, x being the apply's input (updated in place, each DOF by the one
// thread that owns it); with c1 = 0, d is not read.  Unused when b is null.
struct ChebFuse {
  const void* b = nullptr;
  const void* inv = nullptr;  // Dinv, one value per DOF (shared by the columns)
  void* dv = nullptr;
  double c1 = 0.0, c2 = 0.0;
};

// operator selection
struct OpCoef {
  double stiff = 1.0, mass = 0.0;
  double pot = 0.0;      // coefficient of the potential mass operator M_V (DevSpace::vpot)
  int constrained = 0;   // honour the Dirichlet elimination
  int Kl = 0, Kdl = 0;   // leading modes used; 0 = all
  int level = 0;         // p-multigrid level the apply belongs to (0: the space itself)
  // a block of ncol vectors in one launch (grid y = column): column c of x
  // and y starts at c ldx, its element results at c ldl.  Only the
  // fp16-split FP32 path takes ncol > 1 (the p-multigrid block V-cycle)
  int ncol = 1;
  long ldx = 0, ldl = 0;
  ChebFuse cheb;  // scatter forms a Chebyshev step instead of y (cheb.b set)
  // applies counted by the per-kernel timing: the space's own full-order operator
  bool timed() const { return Kl == 0 && level == 0; }
};

// scalar block of the CG loops (device memory)
struct CgScalars {
  double pq, rz, rz_new, alpha, beta, l1, l2, b2;
  double tmp[8];
};

void launch_gather(const DevSpace& S, const double* x, double* xl, int constrained, hipStream_t st);
template <typename T>
void launch_apply(const DevSpace& S, const OpCoef& op, const T* x, T* y, T* yl, double* partial, hipStream_t st);
template <typename T>
void launch_apply_mfma(const DevSpace& S, const OpCoef& op, const T* x, T* yl, hipStream_t st);
bool mfma_supported(const DevSpace& S);
// the MFMA stiffness kernel with its gather fused in (k_elem_gemm_gx)
bool mfma_gx_supported(const DevSpace& S);
// fp16-split path for FP32 vectors (x: FP32 DOF vector, yl: element results,
// element-major, or the entity slab when slab is set)
void launch_apply_mfma_h(const DevSpace& S, const OpCoef& op, const float* x, float* yl, bool slab, hipStream_t st);
// entity-slab element results on (process wide; default on, TSEM_SLAB=0 off)
void set_slab(bool on);
bool slab_enabled();
bool mfma_h_supported(const DevSpace& S);
size_t gemm_h_lds(const DevSpace& S);
size_t gemm_h_colmax_offset(const DevSpace& S);  // DevSpace::hlds_v
// LDS byte offset of the slab tables behind the [K][NE + 1] FP32 Y tile
__host__ __device__ inline size_t slab_tab_offset(int K, int NE) { return ((size_t)K * (NE + 1) * 4 + 15) & ~(size_t)15; }
// bytes of the slab tables staged there: entity offsets and orientation
// words [NE][17], FP32 weights [nrow][JT], 16-bit Y offsets [nrow][JT],
// 16-bit term counts [nrow], rounded up to 4 bytes
__host__ __device__ inline size_t slab_tab_bytes(const DevSpace& S, int NE) {
  const size_t nrj = (size_t)S.sl_nrow * S.sl_JT;
  return ((size_t)NE * 68 + nrj * 6 + (size_t)S.sl_nrow * 2 + 3) & ~(size_t)3;
}
// the slab tables fit between the Y tile and the maxima
bool slab_h_fits(const DevSpace& S);
// the slab scatter takes every chunk span in one pass
bool slab_scatter_fits(const DevSpace& S);
// per-kernel HIP-event timing of full-order applies (kinds below); timing_read
// fills {total ms, launches} per kind
enum { TK_GATHER64 = 0, TK_ELEM64, TK_SCATTER64, TK_GATHER32, TK_ELEM32, TK_SCATTER32, TK_KINDS };
void timing_enable(bool on);
void timing_read(double* out);
int timing_begin(int kind, hipStream_t st);  // -1 when timing is off
void timing_end(int handle, hipStream_t st);
// 0: auto (MFMA when supported), 1: force the scalar element kernel
void set_apply_path(int path);
int apply_path();
// the FP32 operator on a block of columns in one launch per kernel
// (OpCoef::ncol): fp16-split path, stiffness only, default apply path
bool block_apply_supported(const DevSpace& S, const OpCoef& op);
// the FP32 operator of op runs on the fp16-split element kernel (else on the
// FP32 MFMA kernel or the scalar one): stiffness (and potential) operators on
// the default path, except orders up to 4 (K <= 35), where the FP32 MFMA
// kernel with its fused gather is faster (and exact FP32); a block of
// columns takes the same kernel as each column alone
bool split_path(const DevSpace& S, const OpCoef& op);
// captured-graph replay of PCG chunks and V-cycles on (tsem_set_graphs)
bool graphs_enabled();
// the eigen solvers' preconditioner as one block V-cycle per iteration
// (tsem_set_block_vcycle)
bool block_vcycle_enabled();
template <typename T>
void launch_scatter(const DevSpace& S, const T* yl, T* y, hipStream_t st);

// deterministic reductions: partial sums per block then a single-block final
int reduce_blocks(int n);
void launch_reduce_final(const double* partial, int nb, int nvals, double* out, hipStream_t st);

// pq_part: the reduce_blocks(n) partials of p.q; alpha = rz / pq is then
// formed inside the update (no separate final reduction); nullptr uses
// sc->alpha
template <typename T>
void launch_cg_update(const DevSpace& S, int n, T* x, T* r, T* z, const T* p, const T* q, CgScalars* sc,
                      double* partial, int precond_jacobi, const double* pq_part, hipStream_t st);
template <typename T>
void launch_p_update(int n, T* p, const T* z, CgScalars* sc, hipStream_t st);
// beta from the reduce_blocks(n) r.z partials (and |r|_1, |r|_2^2 recorded),
// then p = beta p - z; sc->rz is left for the next pq-partials CG update
template <typename T>
void launch_p_update_beta(int n, T* p, const T* z, CgScalars* sc, const double* partial, hipStream_t st);
void launch_scalar_alpha(CgScalars* sc, hipStream_t st);
void launch_final_alpha(const double* partial, int nb, CgScalars* sc, hipStream_t st);
void launch_final_beta(const double* partial, int nb, int with_rz, CgScalars* sc, hipStream_t st);
void launch_scalar_beta(CgScalars* sc, hipStream_t st);

template <typename T>
void launch_axpby(int n, T a, const T* x, T b, T* y, hipStream_t st);
template <typename T>
void launch_zero(int n, T* y, hipStream_t st);  // y = 0 by a kernel (capturable)
bool timing_on();
template <typename T>
void launch_dot(int n, const T* a, const T* b, double* partial, hipStream_t st);
void launch_cast_f64_f32(int n, const double* x, double scale, float* y, hipStream_t st);
void launch_add_f32_f64(int n, const float* x, double scale, double* y, hipStream_t st);
// y = scale x
void launch_cast_f32_f64(int n, const float* x, double scale, double* y, hipStream_t st);
// partial[b] of |r32 + s r64|_2^2 over reduce_blocks(n) blocks
void launch_gap(int n, const float* r32, const double* r64, double s, double* partial, hipStream_t st);
void launch_residual(const DevSpace& S, int n, const double* b, const double* Ax, double* r, double* partial, hipStream_t st);

// RHS: bl[e][i] = 6V/6 * sum_q w_q f[e][q] phi_q[i] for elements [e0, e1)
// part: scratch for the point-range partial sums (rhs_splits(S, e1 - e0, nq)
// x (e1 - e0) x K doubles), or null for one range
int rhs_splits(const DevSpace& S, int ne, int nq);
void launch_rhs_local(const DevSpace& S, int e0, int e1, int nq, const double* wphi /*nq x K, w folded*/,
                      const double* fq, double* bl, double* part, hipStream_t st);
void launch_qp_coords(const DevSpace& S, const double* xyz_elem /*nt x 12*/, int e0, int e1, int nq,
                      const double* pts /*nq x 3*/, double* out, hipStream_t st);

// ---- block vectors (eigen solvers); column j of a block starts at j * ld
int gram_blocks(int n);  // partial buffer: gram_blocks(n) * p * q doubles
// out[i q + j] = sum_r U[r][i] V[r][j] (p, q <= 96), fixed-order sums
void launch_gram(int n, int p, int q, const double* U, long ldu, const double* V, long ldv, double* partial,
                 double* out, hipStream_t st);
// Y = alpha S C + beta Y with C row-major p x m on device (p <= 96, m <= 32)
void launch_combine(int n, int p, int m, const double* S, long lds, const double* C, double* Y, long ldy,
                    double alpha, double beta, hipStream_t st);
// Y = S C where the output columns from jsplit on take only rows ksplit.. of
// C (the rows above are zero); jsplit a multiple of 16, ksplit of 4
void launch_combine_split(int n, int p, int m, const double* S, long lds, const double* C, double* Y, long ldy,
                          int jsplit, int ksplit, hipStream_t st);
// per-column scaled casts of blocks (ncol columns n apart, scales on the device)
void launch_cast_cols(int n, int ncol, const double* X, const double* s, float* Y, hipStream_t st);
void launch_widen_cols(int n, int ncol, const float* Z, const double* s, double* W, hipStream_t st);
// R[:, j] = AX[:, j] - lam[j] BX[:, j] (lam on device), zero on eliminated rows
void launch_block_resid(int n, int m, const double* AX, const double* BX, const double* lam, const uint8_t* cmask,
                        double* R, hipStream_t st);
// zero eliminated rows of an n x m block; scale rows by inv when not null
void launch_block_mask(int n, int m, const uint8_t* cmask, const double* inv, double* X, hipStream_t st);


// ---- quadrature-point work (values, density, integrals, LDA)
// out[(e - e0) nq + q] = sum_i phi_i(x_q) xl[e][i] for e in [e0, e1)
void launch_qp_values(const DevSpace& S, int e0, int e1, const double* xl, double* out, hipStream_t st);
// rho[e][q] = sum_k occ[k] (sum_i phi_i(x_q) xl_k[e][i])^2, xl_k = xl + k nt K
void launch_density(const DevSpace& S, int nk, const double* xl, const double* occ_dev, double* rho, hipStream_t st);
int integrate_blocks(const DevSpace& S);
// out[v] = sum_e vol_e sum_q w_q f[e][q] m_v(x_q), m = {1, x, y, z, xx, xy, xz, yy, yz, zz}
// (v < nmom; with g != null, f g instead of f), fixed-order sums
void launch_integrate(const DevSpace& S, const double* f, const double* g, int nmom, double* partial, double* out,
                      hipStream_t st);
// out = sum_{e in [e0, e1)} vol_e sum_q w_q (u_h(x_q) - u[(e - e0) nq + q])^2
void launch_l2_chunk(const DevSpace& S, int e0, int e1, const double* xl, const double* u, double* partial,
                     double* out, hipStream_t st);
// gradient error over elements [e0, e1), g: exact gradient at the points ((e1-e0) x nq x 3)
void launch_h1_chunk(const DevSpace& S, int e0, int e1, const double* xl, const double* g, double* partial,
                     double* out, hipStream_t st);
// out[p] = sum_i phi[p][i] xl[ep[p]][i] (NaN where ep[p] < 0): values at
// located points, phi the basis there (n x K)
void launch_point_values(int K, long n, const int32_t* ep, const double* phi, const double* xl, double* out,
                         hipStream_t st);
// LDA, Slater exchange + Perdew-Zunger correlation (unpolarized): potential
// v_xc and energy density e_xc = rho (eps_x + eps_c) per point
void launch_lda_pz(long n, const double* rho, double* vxc, double* exc, hipStream_t st);

// assembled stiffness diagonal and its inverses from the scatter map (grad2: 9 x K x K)
void launch_diagonal(const DevSpace& S, const double* grad2, double* diag, double* inv, float* inv32, hipStream_t st);

}  // namespace tsem
