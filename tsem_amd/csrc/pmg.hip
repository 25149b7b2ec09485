// p-multigrid V-cycle on compact levels with Chebyshev smoothing (pmg.hpp).
#include "pmg.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <random>
#include <stdexcept>
#include <string>

#include "handle.hpp"

namespace tsem {

namespace {
constexpr int BLK = 256;
inline int nblk(long n) { return (int)((n + BLK - 1) / BLK); }

// Chebyshev step: d = c1 d + c2 Dinv (b - Ax); x += d.  Ax null: x = 0 before
// the step (b - Ax = b), and with c1 = 0 neither d nor x is read
// (every kernel here takes column blockIdx.y of a block of vectors, n apart)
template <typename T>
__global__ void k_cheb(int n, const T* __restrict__ b, const T* __restrict__ Ax, const T* __restrict__ inv, T c1,
                       T c2, T* __restrict__ dv, T* __restrict__ x) {
  const int i = blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  const size_t o = (size_t)blockIdx.y * n;
  b += o;
  if (Ax) Ax += o;
  dv += o;
  x += o;
  const T r = Ax ? b[i] - Ax[i] : b[i];
  T dn = c2 * inv[i] * r;
  if (c1 != T(0)) dn += c1 * dv[i];
  dv[i] = dn;
  x[i] = (c1 != T(0) || Ax) ? x[i] + dn : dn;
}

// coarse right-hand side: bc[i] = bf[m[i]] - Axf[m[i]] (restricted residual)
template <typename T>
__global__ void k_restrict_res(int nc, long nf, const int32_t* __restrict__ m, const T* __restrict__ bf,
                               const T* __restrict__ axf, T* __restrict__ bc) {
  const int i = blockIdx.x * BLK + threadIdx.x;
  const size_t of = (size_t)blockIdx.y * nf, oc = (size_t)blockIdx.y * nc;
  if (i < nc) bc[oc + i] = bf[of + m[i]] - axf[of + m[i]];
}

// xf[m[i]] += xc[i] (zero extension of the coarse correction; m injective)
template <typename T>
__global__ void k_prolong_add(int nc, long nf, const int32_t* __restrict__ m, const T* __restrict__ xc,
                              T* __restrict__ xf) {
  const int i = blockIdx.x * BLK + threadIdx.x;
  const size_t of = (size_t)blockIdx.y * nf, oc = (size_t)blockIdx.y * nc;
  if (i < nc) xf[of + m[i]] += xc[oc + i];
}

// power iteration for lambda_max(D^-1 A) in the D inner product: partials of
// v.w, v.D v and w.D^-1 w with w = A v (three reduce_blocks(n) segments)
__global__ void __launch_bounds__(BLK) k_pow_partials(int n, const double* __restrict__ v, const double* __restrict__ w,
                                                      const double* __restrict__ dg, const double* __restrict__ inv,
                                                      double* __restrict__ part) {
  __shared__ double red[3][BLK / 64];
  double a = 0, b = 0, c = 0;
  for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK) {
    a += v[i] * w[i];
    b += v[i] * v[i] * dg[i];
    c += w[i] * w[i] * inv[i];
  }
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_xor(a, o, 64);
    b += __shfl_xor(b, o, 64);
    c += __shfl_xor(c, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = a;
    red[1][threadIdx.x >> 6] = b;
    red[2][threadIdx.x >> 6] = c;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    double s = 0;
    for (int k = 0; k < BLK / 64; ++k) s += red[threadIdx.x][k];
    part[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = s;
  }
}
// v = D^-1 w / sqrt(w.D^-1 w) (the next iterate, unit in the D norm; sums[2] on the device)
__global__ void k_pow_next(int n, const double* __restrict__ w, const double* __restrict__ inv,
                           const double* __restrict__ sums, double* __restrict__ v) {
  const int i = blockIdx.x * BLK + threadIdx.x;
  const double s = sums[2] > 0 ? 1.0 / sqrt(sums[2]) : 0.0;
  if (i < n) v[i] = inv[i] * w[i] * s;
}

// y = A x for a CSR matrix, one row per thread, entries in column order
template <typename T>
__global__ void k_csr_spmv(int n, const int32_t* __restrict__ rs, const int32_t* __restrict__ col,
                           const T* __restrict__ val, const T* __restrict__ x, T* __restrict__ y) {
  const int i = blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  const size_t o = (size_t)blockIdx.y * n;
  T acc = T(0);
  for (int k = rs[i]; k < rs[i + 1]; ++k) acc += val[k] * x[o + col[k]];
  y[o + i] = acc;
}

// one Chebyshev step on the assembled level, with its product fused in:
// d = c1 d + c2 Dinv (b - A xin), xout = xin + d (xin and xout distinct: the
// rows read their neighbours' xin)
// one Chebyshev step with the CSR product fused in, CSR_LPR lanes per row:
// lane q of a row's group sums the row's entries q, q + CSR_LPR, ..., and
// the group adds its partial sums by a fixed xor-shuffle tree (the order is
// fixed, so the step is reproducible).  A row per lane left half the SIMDs
// idle on the 36k-row order-1 level of the 32^3 mesh.
constexpr int CSR_LPR = 4;
template <typename T>
__global__ void k_cheb_csr(int n, const int32_t* __restrict__ rs, const int32_t* __restrict__ col,
                           const T* __restrict__ val, const T* __restrict__ b, const T* __restrict__ xin,
                           const T* __restrict__ inv, T c1, T c2, T* __restrict__ dv, T* __restrict__ xout) {
  const int gt = blockIdx.x * BLK + threadIdx.x;
  const int i = gt / CSR_LPR, q = gt % CSR_LPR;
  const size_t o = (size_t)blockIdx.y * n;
  T acc = T(0);
  if (i < n)
    for (int k = rs[i] + q; k < rs[i + 1]; k += CSR_LPR) acc += val[k] * xin[o + col[k]];
  // every lane of the block reaches the shuffles (BLK is a multiple of CSR_LPR)
#pragma unroll
  for (int d = CSR_LPR / 2; d >= 1; d /= 2) acc += __shfl_xor(acc, d, CSR_LPR);
  if (i >= n || q != 0) return;
  const T r = b[o + i] - acc;
  T dn = c2 * inv[i] * r;
  if (c1 != T(0)) dn += c1 * dv[o + i];
  dv[o + i] = dn;
  xout[o + i] = xin[o + i] + dn;
}

template <typename T>
T* as(void* p) { return static_cast<T*>(p); }
}  // namespace

// The level-l operator assembled from its space's element matrices and
// gather terms (sum_e G_e^T K_e G_e), with the elimination of the level's
// mask: an eliminated row or column keeps only its diagonal entry.
void Pmg::assemble_level(int l, hipStream_t st) {
  const tsem_space* c = sub[l - 1];
  const Space& S = c->h;
  const int K = S.K;
  struct Tri { int32_t r, c; double v; };
  std::vector<Tri> tri;
  tri.reserve((size_t)S.topo.nt * K * K);
  std::vector<std::pair<int64_t, double>> terms;
  std::vector<double> Ke;
  std::vector<std::vector<std::pair<int64_t, double>>> G(K);
  for (int e = 0; e < S.topo.nt; ++e) {
    S.element_matrix(e, 1.0, 0.0, Ke);
    for (int i = 0; i < K; ++i) {
      G[i].clear();
      S.gather_terms(e, i, G[i]);
    }
    for (int i = 0; i < K; ++i)
      for (int j = 0; j < K; ++j) {
        const double kij = Ke[(size_t)i * K + j];
        if (kij == 0.0) continue;
        for (auto& a : G[i])
          for (auto& b : G[j]) tri.push_back({(int32_t)a.first, (int32_t)b.first, a.second * kij * b.second});
      }
  }
  std::sort(tri.begin(), tri.end(), [](const Tri& a, const Tri& b) { return a.r < b.r || (a.r == b.r && a.c < b.c); });
  const int64_t n = S.ndof;
  const auto& m = c->cmask;
  std::vector<int32_t> rs(n + 1, 0), col;
  std::vector<double> v;
  col.reserve(tri.size() / 4);
  v.reserve(tri.size() / 4);
  for (size_t k = 0; k < tri.size();) {
    size_t k2 = k;
    double sum = 0.0;
    while (k2 < tri.size() && tri[k2].r == tri[k].r && tri[k2].c == tri[k].c) sum += tri[k2++].v;
    const int32_t r = tri[k].r, cc = tri[k].c;
    if (r == cc || (!m[r] && !m[cc])) {
      col.push_back(cc);
      v.push_back(sum);
      rs[r + 1]++;
    }
    k = k2;
  }
  for (int64_t i = 0; i < n; ++i) rs[i + 1] += rs[i];
  HIPCHK(hipMalloc(&csr_rs, (n + 1) * 4));
  HIPCHK(hipMalloc(&csr_col, col.size() * 4 + 4));
  HIPCHK(hipMalloc(&csr_v64, v.size() * 8 + 8));
  HIPCHK(hipMalloc(&csr_v32, v.size() * 4 + 4));
  std::vector<float> v32(v.begin(), v.end());
  HIPCHK(hipMemcpyAsync(csr_rs, rs.data(), (n + 1) * 4, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(csr_col, col.data(), col.size() * 4, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(csr_v64, v.data(), v.size() * 8, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(csr_v32, v32.data(), v32.size() * 4, hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
  csr_level = l;
}

const DevSpace& Pmg::dev(int l) const { return l == 0 ? fine->d : sub[l - 1]->d; }

Pmg::Pmg(tsem_space* fsp, hipStream_t st) : fine(fsp) {
  const Space& S = fine->h;
  const int p = S.M;
  // levels: p, p / 2 (when above 1), 1 -- measured faster than the
  // reference's p, p - 1, ..., 1 (n_project = p - 1) on the 32^3, p = 6 mesh:
  // the orders between cost a full pass over the elements each and save
  // fewer iterations than they cost; TSEM_PMG_ORDERS lists other coarse
  // orders (descending, below p; "5,4,3,2,1" is the reference's set)
  order.push_back(p);
  if (const char* env = getenv("TSEM_PMG_ORDERS")) {
    for (const char* q = env; *q;) {
      char* e = nullptr;
      const long o = strtol(q, &e, 10);
      if (e == q) break;
      if (o >= 1 && o < order.back()) order.push_back((int)o);
      q = (*e == ',') ? e + 1 : e;
    }
  } else {
    if (p / 2 > 1) order.push_back(p / 2);
    if (p > 1) order.push_back(1);
  }
  if (const char* env = getenv("TSEM_PMG_COARSE")) coarse_deg = std::max(1, atoi(env));
  if (const char* env = getenv("TSEM_PMG_CRATIO")) coarse_ratio = std::max(2.0, atof(env));
  const int L = (int)order.size();
  ndof_level.assign(L, 0);
  ndof_level[0] = S.ndof;
  sub.assign(L - 1, nullptr);
  map.assign(L, nullptr);
  vx.assign(L, nullptr); vb.assign(L, nullptr); vr.assign(L, nullptr); vd.assign(L, nullptr);
  try {
    // position of every fine DOF in each level (-1 when its degree is above
    // the level's order): the rank among the fine DOFs of degree <= p_l
    const int64_t N = S.ndof;
    std::vector<int64_t> prev_pos(N);
    for (int64_t i = 0; i < N; ++i) prev_pos[i] = i;
    for (int l = 1; l < L; ++l) {
      const int pl = order[l];
      tsem_space* c = nullptr;
      int rc = tsem_space_create(fine->ctx, pl, S.topo.nv, S.xyz.data(), S.topo.nt, S.tet.data(), S.femdof.data(), 1,
                                 &c);
      if (rc != TSEM_OK) throw std::runtime_error(std::string("p-multigrid level: ") + tsem_last_error());
      sub[l - 1] = c;
      const int64_t Nl = c->h.ndof;
      ndof_level[l] = Nl;
      std::vector<int64_t> pos(N, -1);
      int64_t k = 0;
      for (int64_t i = 0; i < N; ++i)
        if (S.sum_dof[i] <= pl) pos[i] = k++;
      if (k != Nl) throw std::logic_error("p-multigrid level size differs from the order-p_l space");
      // the order-p_l space must lay its geometries out as the level does:
      // each geometry's first DOF at its level position
      {
        const int64_t cnt[4] = {S.topo.nv, S.topo.ne, S.topo.nf, S.topo.nt};
        if (c->h.dof_start.size() != S.dof_start.size())
          throw std::logic_error("p-multigrid level has another geometry count");
        size_t g = 0;
        for (int dim = 0; dim < 4; ++dim)
          for (int64_t j = 0; j < cnt[dim]; ++j, ++g)
            if (c->h.ndg[dim] > 0 && pos[S.dof_start[g]] != c->h.dof_start[g])
              throw std::logic_error("p-multigrid level layout differs from the order-p_l space");
      }
      // injection map into the previous level, and the level's elimination
      std::vector<int32_t> m(Nl);
      for (int64_t i = 0; i < N; ++i)
        if (pos[i] >= 0) m[pos[i]] = (int32_t)prev_pos[i];
      for (int64_t i = 0; i < N; ++i)
        if (pos[i] >= 0) c->cmask[pos[i]] = fine->cmask[i];
      upload_mask(c, st);
      HIPCHK(hipMalloc(&map[l], (size_t)Nl * 4));
      HIPCHK(hipMemcpyAsync(map[l], m.data(), (size_t)Nl * 4, hipMemcpyHostToDevice, st));
      prev_pos.swap(pos);
      HIPCHK(hipStreamSynchronize(st));
    }
    if (L > 1 && order[L - 1] == 1) assemble_level(L - 1, st);
    for (int l = 0; l < L; ++l) {
      const size_t b = (size_t)ndof_level[l] * 8;
      HIPCHK(hipMalloc(&vr[l], b));
      HIPCHK(hipMalloc(&vd[l], b));
      if (l > 0) {
        HIPCHK(hipMalloc(&vx[l], b));
        HIPCHK(hipMalloc(&vb[l], b));
      }
    }
    // lambda_max(D^-1 A_l) by power iteration on the device, one read at the end
    lmax.assign(L, 2.0);
    constexpr int ITS = 15;
    double* part = nullptr;
    double* sums = nullptr;
    HIPCHK(hipMalloc(&part, (size_t)3 * reduce_blocks((int)N) * 8));
    HIPCHK(hipMalloc(&sums, (size_t)4 * L * 8));
    std::mt19937_64 rng(12345);
    std::uniform_real_distribution<double> U(-1, 1);
    // start vectors of every level, alive until the final synchronization
    size_t tot = 0;
    for (int l = 0; l < L; ++l) tot += (size_t)ndof_level[l];
    std::vector<double> v0(tot);
    for (double& a : v0) a = U(rng);
    size_t off = 0;
    for (int l = 0; l < L; ++l) {
      const int n = (int)ndof_level[l];
      const DevSpace& d = dev(l);
      double* v = as<double>(vd[l]);
      double* w = as<double>(vr[l]);
      HIPCHK(hipMemcpyAsync(v, v0.data() + off, (size_t)n * 8, hipMemcpyHostToDevice, st));
      off += (size_t)n;
      const int nb = reduce_blocks(n);
      for (int it = 0; it < ITS; ++it) {
        apply_level<double>(l, v, w, st);
        hipLaunchKernelGGL(k_pow_partials, dim3(nb), dim3(BLK), 0, st, n, (const double*)v, (const double*)w, d.diag,
                           d.invdiag, part);
        launch_reduce_final(part, nb, 3, sums + 4 * l, st);
        hipLaunchKernelGGL(k_pow_next, dim3(nblk(n)), dim3(BLK), 0, st, n, (const double*)w, d.invdiag,
                           (const double*)(sums + 4 * l), v);
      }
    }
    std::vector<double> h(4 * (size_t)L);
    HIPCHK(hipMemcpyAsync(h.data(), sums, h.size() * 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (int l = 0; l < L; ++l) lmax[l] = (h[4 * l + 1] > 0) ? h[4 * l] / h[4 * l + 1] : 2.0;
    (void)hipFree(part);
    (void)hipFree(sums);
  } catch (...) {
    release();
    throw;
  }
}

Pmg::~Pmg() { release(); }

void Pmg::release() {
  if (vexec) (void)hipGraphExecDestroy(vexec);
  vexec = nullptr;
  vkey.clear();
  for (auto* v : {&bx, &bb, &br, &bd, &byl})
    for (void*& p : *v) {
      if (p) (void)hipFree(p);
      p = nullptr;
    }
  bcap = 0;
  for (auto* v : {&vx, &vb, &vr, &vd})
    for (void*& p : *v) {
      if (p) (void)hipFree(p);
      p = nullptr;
    }
  for (int32_t*& p : map) {
    if (p) (void)hipFree(p);
    p = nullptr;
  }
  for (void* p : {(void*)csr_rs, (void*)csr_col, (void*)csr_v64, (void*)csr_v32})
    if (p) (void)hipFree(p);
  csr_rs = nullptr; csr_col = nullptr; csr_v64 = nullptr; csr_v32 = nullptr;
  csr_level = -1;
  for (tsem_space*& c : sub) {
    if (c) (void)tsem_space_destroy(c);
    c = nullptr;
  }
}

long Pmg::yl_cols(int l) const {
  const tsem_space* s = l == 0 ? fine : sub[l - 1];
  return (long)(s->yl32.n / sizeof(float));
}

bool Pmg::block_supported() const {
  for (int l = 0; l < levels(); ++l) {
    if (l == csr_level) continue;
    OpCoef op;
    op.constrained = 1;
    op.level = l;
    op.ncol = 2;
    if (!block_apply_supported(dev(l), op)) return false;
  }
  return true;
}

void Pmg::block_reserve(int ncol) {
  if (ncol <= bcap) return;
  // the captured V-cycle may hold the buffers about to be replaced
  if (vexec) (void)hipGraphExecDestroy(vexec);
  vexec = nullptr;
  vkey.clear();
  const int L = levels();
  for (auto* v : {&bx, &bb, &br, &bd, &byl}) {
    for (void*& p : *v)
      if (p) (void)hipFree(p);
    v->assign(L, nullptr);
  }
  bcap = 0;
  for (int l = 0; l < L; ++l) {
    const size_t b = (size_t)ncol * ndof_level[l] * sizeof(float);
    HIPCHK(hipMalloc(&br[l], b));
    HIPCHK(hipMalloc(&bd[l], b));
    if (l > 0) {
      HIPCHK(hipMalloc(&bx[l], b));
      HIPCHK(hipMalloc(&bb[l], b));
    }
    if (l != csr_level) HIPCHK(hipMalloc(&byl[l], (size_t)ncol * yl_cols(l) * sizeof(float)));
  }
  bcap = ncol;
}

template <typename T>
void Pmg::apply_level(int l, const T* x, T* y, hipStream_t st, int ncol, const ChebFuse& cf) {
  const int n = (int)ndof_level[l];
  if (l == csr_level) {
    if (cf.b) throw std::logic_error("the assembled level takes no fused Chebyshev step");
    const T* val = sizeof(T) == 8 ? (const T*)csr_v64 : (const T*)csr_v32;
    hipLaunchKernelGGL(k_csr_spmv<T>, dim3(nblk(n), ncol), dim3(BLK), 0, st, n, (const int32_t*)csr_rs,
                       (const int32_t*)csr_col, val, x, y);
    return;
  }
  OpCoef op;
  op.constrained = 1;
  op.level = l;
  op.cheb = cf;
  tsem_space* s = l == 0 ? fine : sub[l - 1];
  T* yl = sizeof(T) == 8 ? (T*)s->yl64.p : (T*)s->yl32.p;
  if (ncol > 1) {
    if (sizeof(T) != 4 || ncol > bcap) throw std::logic_error("block V-cycle: FP32 columns within the reserve");
    op.ncol = ncol;
    op.ldx = n;
    op.ldl = yl_cols(l);
    yl = as<T>(byl[l]);
  }
  launch_apply<T>(s->d, op, x, y, yl, nullptr, st);
}

// Chebyshev iteration of the given degree on D^-1 A_l x = D^-1 b over
// [hi / 30, hi], hi = 1.1 lmax (the power iteration approaches lmax from below)
template <typename T>
void Pmg::smooth_level(int l, T* x, const T* b, int degree, bool zero_init, hipStream_t st, int ncol) {
  const int n = (int)ndof_level[l];
  const DevSpace& d = dev(l);
  T* Ax = as<T>(ncol > 1 ? br[l] : vr[l]);
  T* dv = as<T>(ncol > 1 ? bd[l] : vd[l]);
  const dim3 grid(nblk(n), ncol);
  const T* inv = (sizeof(T) == 8) ? (const T*)d.invdiag : (const T*)d.invdiag32;
  const double hi = 1.1 * lmax[l], lo = hi / (l + 1 == levels() && l > 0 ? coarse_ratio : 30.0);
  const double theta = (hi + lo) / 2, delta = (hi - lo) / 2, sigma = theta / delta;
  double rho_old = 1.0 / sigma;
  if (l == csr_level && zero_init) {
    // the assembled coarsest level: each later step is one launch, its
    // product fused in, alternating between x and the product buffer (free
    // here); the first step writes the buffer that the last step leaves in x
    T* buf[2] = {x, Ax};
    int cur = (degree - 1) % 2;
    hipLaunchKernelGGL(k_cheb<T>, grid, dim3(BLK), 0, st, n, b, (const T*)nullptr, inv, T(0), T(1.0 / theta), dv,
                       buf[cur]);
    const T* val = sizeof(T) == 8 ? (const T*)csr_v64 : (const T*)csr_v32;
    const dim3 cgrid(nblk(n * CSR_LPR), ncol);
    for (int k = 1; k < degree; ++k) {
      const double rho = 1.0 / (2 * sigma - rho_old);
      hipLaunchKernelGGL(k_cheb_csr<T>, cgrid, dim3(BLK), 0, st, n, (const int32_t*)csr_rs, (const int32_t*)csr_col,
                         val, b, (const T*)buf[cur], inv, T(rho * rho_old), T(2 * rho / delta), dv, buf[1 - cur]);
      cur = 1 - cur;
      rho_old = rho;
    }
    return;
  }
  // steps after a product: the step is formed by the apply's scatter
  // (ChebFuse), with no A x vector and no separate pass over the vectors
  auto fused = [&](double c1, double c2) {
    ChebFuse cf;
    cf.b = b;
    cf.inv = inv;
    cf.dv = dv;
    cf.c1 = c1;
    cf.c2 = c2;
    apply_level<T>(l, x, x, st, ncol, cf);
  };
  if (zero_init)
    hipLaunchKernelGGL(k_cheb<T>, grid, dim3(BLK), 0, st, n, b, (const T*)nullptr, inv, T(0), T(1.0 / theta), dv, x);
  else
    fused(0.0, 1.0 / theta);
  for (int k = 1; k < degree; ++k) {
    const double rho = 1.0 / (2 * sigma - rho_old);
    fused(rho * rho_old, 2 * rho / delta);
    rho_old = rho;
  }
}

template <typename T>
void Pmg::vcycle(const T* r, T* z, int smooth, hipStream_t st, int ncol) {
  const int L = levels();
  if (smooth < 1) smooth = 1;
  if (ncol < 1) throw std::invalid_argument("V-cycle needs at least one column");
  if (ncol > 1) {
    if (sizeof(T) != 4 || !block_supported()) throw std::logic_error("block V-cycle needs the FP32 block operator");
    block_reserve(ncol);
  }
  std::vector<T*> X(L), B(L);
  X[0] = z;
  B[0] = const_cast<T*>(r);
  for (int l = 1; l < L; ++l) {
    X[l] = as<T>(ncol > 1 ? bx[l] : vx[l]);
    B[l] = as<T>(ncol > 1 ? bb[l] : vb[l]);
  }
  for (int l = 0; l + 1 < L; ++l) {
    smooth_level<T>(l, X[l], B[l], smooth, true, st, ncol);
    T* Ax = as<T>(ncol > 1 ? br[l] : vr[l]);
    apply_level<T>(l, X[l], Ax, st, ncol);
    const int nc = (int)ndof_level[l + 1];
    hipLaunchKernelGGL(k_restrict_res<T>, dim3(nblk(nc), ncol), dim3(BLK), 0, st, nc, ndof_level[l],
                       (const int32_t*)map[l + 1], (const T*)B[l], (const T*)Ax, B[l + 1]);
  }
  smooth_level<T>(L - 1, X[L - 1], B[L - 1], L > 1 ? coarse_deg : smooth, true, st, ncol);
  for (int l = L - 2; l >= 0; --l) {
    const int nc = (int)ndof_level[l + 1];
    hipLaunchKernelGGL(k_prolong_add<T>, dim3(nblk(nc), ncol), dim3(BLK), 0, st, nc, ndof_level[l],
                       (const int32_t*)map[l + 1], (const T*)X[l + 1], X[l]);
    smooth_level<T>(l, X[l], B[l], smooth, false, st, ncol);
  }
}

template void Pmg::vcycle<double>(const double*, double*, int, hipStream_t, int);
template void Pmg::vcycle<float>(const float*, float*, int, hipStream_t, int);

template <typename T>
void Pmg::vcycle_replay(const T* r, T* z, int smooth, hipStream_t st, int ncol) {
  const std::vector<const void*> key = {r, z, (const void*)(intptr_t)smooth, (const void*)sizeof(T),
                                        (const void*)(intptr_t)ncol, (const void*)(intptr_t)apply_path(),
                                        (const void*)(intptr_t)slab_enabled()};
  if (!graphs_enabled() || timing_on() || key == vfailed) {
    vcycle<T>(r, z, smooth, st, ncol);
    return;
  }
  if (ncol > 1) block_reserve(ncol);  // no allocation inside the capture
  if (!vexec || vkey != key) {
    if (vexec) (void)hipGraphExecDestroy(vexec);
    vexec = nullptr;
    vkey.clear();
    hipGraph_t g = nullptr;
    bool ok = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) == hipSuccess;
    if (ok) {
      try {
        vcycle<T>(r, z, smooth, st, ncol);
      } catch (...) {
        ok = false;
      }
      if (hipStreamEndCapture(st, &g) != hipSuccess) ok = false;
      if (ok && g && hipGraphInstantiate(&vexec, g, nullptr, nullptr, 0) != hipSuccess) {
        vexec = nullptr;
        ok = false;
      }
      if (g) (void)hipGraphDestroy(g);
    }
    if (!ok || !vexec) {
      // the captured launches never ran: this V-cycle runs eagerly, and so
      // does every later one with this key
      if (vexec) (void)hipGraphExecDestroy(vexec);
      vexec = nullptr;
      vfailed = key;
      (void)hipGetLastError();
      vcycle<T>(r, z, smooth, st, ncol);
      return;
    }
    vkey = key;
  }
  HIPCHK(hipGraphLaunch(vexec, st));
}

template void Pmg::vcycle_replay<double>(const double*, double*, int, hipStream_t, int);
template void Pmg::vcycle_replay<float>(const float*, float*, int, hipStream_t, int);

}  // namespace tsem
