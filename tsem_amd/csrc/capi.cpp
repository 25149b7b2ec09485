// C ABI of the TSEM library (include/tsem.h): contexts, spaces, device
// buffers, boundary conditions and the PCG / mixed-precision drivers.
#include "../../include/tsem.h"

#include <hip/hip_runtime.h>

#include <dlfcn.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cmath>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "handle.hpp"

using namespace tsem;

namespace tsem {
thread_local std::string g_err;
}

// PCG iterations replayed as captured graphs (process wide; tsem_set_graphs)
static bool g_pcg_graphs = [] {
  const char* v = getenv("TSEM_PCG_GRAPH");
  return !(v && v[0] == '0');
}();

// alpha formed inside the CG update kernel rather than by its own
// single-block reduction (TSEM_FUSED_ALPHA=0 keeps the separate kernel; the
// two give the same alpha bitwise)
static const bool g_fused_alpha = [] {
  const char* v = getenv("TSEM_FUSED_ALPHA");
  return !(v && v[0] == '0');
}();
// beta formed inside the p update likewise (needs the in-update alpha, which
// moves the latest r.z into place; TSEM_FUSED_BETA=0 keeps k_final_beta)
static const bool g_fused_beta = [] {
  const char* v = getenv("TSEM_FUSED_BETA");
  return !(v && v[0] == '0');
}();

const char* tsem_last_error(void) { return g_err.c_str(); }

// Cooperative solve classes (process wide): while any foreground solve runs,
// a background PCG solve waits between its chunks of iterations, each time
// for at most as long as its previous chunk took: short foreground requests
// get the device mostly alone, and a background solve keeps at least about
// half of it.
namespace {
std::mutex g_fg_mu;
std::condition_variable g_fg_cv;
int g_fg_active = 0;

struct ForegroundScope {
  bool on;
  explicit ForegroundScope(const tsem_ctx* c) : on(c && c->cls == 1) {
    if (!on) return;
    std::lock_guard<std::mutex> lk(g_fg_mu);
    ++g_fg_active;
  }
  ~ForegroundScope() {
    if (!on) return;
    {
      std::lock_guard<std::mutex> lk(g_fg_mu);
      --g_fg_active;
    }
    g_fg_cv.notify_all();
  }
};

// max_seconds <= 0: until no foreground work is active
// how long a background wait may last, as a multiple of the caller's
// previous chunk (TSEM_BG_YIELD, default 1: background work keeps at least
// half the device while foreground work runs)
static const double g_bg_yield = [] {
  const char* v = getenv("TSEM_BG_YIELD");
  const double f = v ? atof(v) : 1.0;
  return f > 0 ? f : 1.0;
}();

void background_wait(const tsem_ctx* c, double max_seconds) {
  if (!c || c->cls != 2) return;
  if (max_seconds > 0) max_seconds *= g_bg_yield;
  std::unique_lock<std::mutex> lk(g_fg_mu);
  if (max_seconds <= 0) g_fg_cv.wait(lk, [] { return g_fg_active == 0; });
  else g_fg_cv.wait_for(lk, std::chrono::duration<double>(max_seconds), [] { return g_fg_active == 0; });
}
}  // namespace


static void ctx_release(tsem_ctx* c) {
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

// the face table ships next to the library: <dir of libtsem_amd.so>/data/face_table.txt
static std::string face_table_path() {
  const char* env = getenv("TSEM_FACE_TABLE");
  if (env) return env;
  Dl_info info;
  if (dladdr((void*)&face_table_path, &info) && info.dli_fname) {
    std::string f(info.dli_fname);
    size_t k = f.find_last_of('/');
    std::string dir = (k == std::string::npos) ? std::string(".") : f.substr(0, k);
    return dir + "/data/face_table.txt";
  }
  return "data/face_table.txt";
}

// Upload the per-DOF elimination mask and derive the per-element geometry
// bits the tiled gather reads (vertex k: bit k, edge k: bit 4+k, face k: bit
// 10+k, interior: bit 14).  Eliminations act on whole geometries, so the
// first DOF of a geometry speaks for all of them; checked here.
void upload_mask(tsem_space* sp, hipStream_t st) {
  const Space& S = sp->h;
  const auto& m = sp->cmask;
  auto geo = [&](int64_t start, int n) -> int {
    if (n <= 0) return 0;
    int v = m[start];
    for (int i = 1; i < n; ++i)
      if (m[start + i] != v) throw std::logic_error("elimination mask is not uniform over a geometry");
    return v;
  };
  std::vector<int32_t> gm(S.desc.size());
  for (size_t e = 0; e < S.desc.size(); ++e) {
    const ElemDesc& D = S.desc[e];
    int b = 0;
    for (int k = 0; k < 4; ++k) b |= geo(D.v[k], S.ndg[0] > 0 ? 1 : 0) << k;
    for (int k = 0; k < 6; ++k) b |= geo(D.e[k] & 0x7fffffff, S.ndg[1]) << (4 + k);
    for (int k = 0; k < 4; ++k) b |= geo(D.f[k], S.ndg[2]) << (10 + k);
    b |= geo(D.t, S.ndg[3]) << 14;
    gm[e] = b;
  }
  upload(sp->cmask_d, sp->cmask, st);
  upload(sp->gmask, gm, st);
  sp->d.cmask = sp->cmask_d.as<uint8_t>();
  sp->d.gmask = sp->gmask.as<int32_t>();
}

extern "C" {

int tsem_ctx_create(int device, tsem_ctx** out) {
  API_BEGIN
  if (!out) throw std::invalid_argument("null output");
  int n = 0;
  HIPCHK(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) throw std::invalid_argument("device index out of range");
  HIPCHK(hipSetDevice(device));
  tsem::scratch_init(device);
  auto* c = new tsem_ctx();
  c->device = device;
  HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  *out = c;
  API_END
}

int tsem_foreground_begin(tsem_ctx* ctx) {
  API_BEGIN
  if (!ctx) throw std::invalid_argument("null context");
  std::lock_guard<std::mutex> lk(g_fg_mu);
  ++g_fg_active;
  API_END
}

int tsem_foreground_end(tsem_ctx* ctx) {
  API_BEGIN
  if (!ctx) throw std::invalid_argument("null context");
  {
    std::lock_guard<std::mutex> lk(g_fg_mu);
    if (g_fg_active <= 0) throw std::logic_error("tsem_foreground_end without tsem_foreground_begin");
    --g_fg_active;
  }
  g_fg_cv.notify_all();
  API_END
}

int tsem_background_wait(tsem_ctx* ctx, double max_seconds) {
  API_BEGIN
  if (!ctx) throw std::invalid_argument("null context");
  background_wait(ctx, max_seconds);
  API_END
}

int tsem_ctx_set_class(tsem_ctx* ctx, int cls) {
  API_BEGIN
  if (!ctx) throw std::invalid_argument("null context");
  if (cls < 0 || cls > 2) throw std::invalid_argument("class must be 0 (normal), 1 (foreground) or 2 (background)");
  ctx->cls = cls;
  API_END
}

int tsem_ctx_destroy(tsem_ctx* ctx) {
  API_BEGIN
  if (!ctx) return TSEM_OK;
  if (ctx->nspaces > 0) ctx->closing = true;
  else ctx_release(ctx);
  API_END
}

int tsem_space_create(tsem_ctx* ctx, int M, int nv, const double* xyz, int nt, const int32_t* tet,
                      const int32_t* femdof_weight, int qorder, tsem_space** out) {
  API_BEGIN
  if (!out || !xyz || !tet) throw std::invalid_argument("null argument");
  if (M < 1 || M > 12) throw std::invalid_argument("polynomial order must be in 1..12");
  if (nv < 4 || nt < 1) throw std::invalid_argument("empty mesh");
  if (qorder < 1 || qorder > 64) throw std::invalid_argument("qorder must be in 1..64");
  if (ctx) HIPCHK(hipSetDevice(ctx->device));
  auto sp = std::make_unique<tsem_space>();
  sp->ctx = ctx;
  Space& S = sp->h;
  try {
    S.build(M, nv, xyz, nt, tet, femdof_weight, face_table_path());
  } catch (const std::runtime_error& e) {
    throw std::invalid_argument(e.what());
  }
  for (int t = 0; t < nt; ++t)
    if (!(S.vol[t] > 0)) throw std::invalid_argument("degenerate tetrahedron");
  // quadrature
  sp->qorder = qorder;
  std::vector<long double> t, w;
  gauss_legendre01(qorder, t, w);
  int q = qorder;
  sp->nq1 = q; sp->nq2 = q * q; sp->nq3 = q * q * q;
  sp->t1.resize(q); sp->w1.resize(q);
  for (int i = 0; i < q; ++i) { sp->t1[i] = (double)t[i]; sp->w1[i] = (double)w[i]; }
  sp->p2.resize(3 * (size_t)q * q); sp->w2.resize((size_t)q * q);
  for (int i = 0; i < q; ++i)
    for (int j = 0; j < q; ++j) {
      size_t k = (size_t)i * q + j;
      long double a = t[i] * (1 - t[j]), b = t[j];
      sp->p2[3 * k] = (double)a; sp->p2[3 * k + 1] = (double)b; sp->p2[3 * k + 2] = (double)(1 - a - b);
      sp->w2[k] = (double)(2 * (1 - t[j]) * w[i] * w[j]);
    }
  sp->p3.resize(3 * (size_t)sp->nq3); sp->w3.resize(sp->nq3);
  for (int i = 0; i < q; ++i)
    for (int j = 0; j < q; ++j)
      for (int l = 0; l < q; ++l) {
        size_t k = ((size_t)i * q + j) * q + l;
        sp->p3[3 * k] = (double)(t[i] * (1 - t[j]) * (1 - t[l]));
        sp->p3[3 * k + 1] = (double)(t[j] * (1 - t[l]));
        sp->p3[3 * k + 2] = (double)t[l];
        sp->w3[k] = (double)(6 * (1 - t[j]) * (1 - t[l]) * (1 - t[l]) * w[i] * w[j] * w[l]);
      }
  sp->cmask.assign(S.ndof, 0);
  tsem::build_slab(S, sp->slab);
  if (!ctx) {  // host-only space: setup and map export without a device
    S.diagonal(1.0, 0.0, sp->diag);
    *out = sp.release();
    return TSEM_OK;
  }
  hipStream_t st = ctx->stream;
  // device data
  DevSpace& d = sp->d;
  d.nt = nt; d.K = S.K; d.Kd = S.ref.Kd; d.M = M; d.nface = S.table.nface; d.ndof = (int)S.ndof;
  upload(sp->desc, S.desc, st);
  std::vector<int32_t> slot(S.K);
  for (int i = 0; i < S.K; ++i) slot[i] = S.slot_kind[i] | (S.slot_geo[i] << 2) | (S.slot_sub[i] << 5);
  upload(sp->slot, slot, st);
  std::vector<int32_t> fptr(1, 0), fkey;
  std::vector<double> fw;
  for (auto& row : S.table.terms) {
    for (auto& tm : row) { fkey.push_back(tm.ie | (tm.idx << 2)); fw.push_back(tm.w); }
    fptr.push_back((int32_t)fkey.size());
  }
  upload(sp->ft_ptr, fptr, st); upload(sp->ft_key, fkey, st); upload(sp->ft_w, fw, st);
  {
    const int nd = S.ndg[2];
    int J = 1;
    for (int v = 0; v < 6; ++v)
      for (int dd = 0; dd < nd; ++dd) J = std::max(J, (int)S.table.terms[(size_t)v * S.table.nface + dd].size());
    std::vector<int32_t> cn((size_t)6 * nd, 0), ck((size_t)6 * nd * J, 0);
    std::vector<double> cw((size_t)6 * nd * J, 0.0);
    for (int v = 0; v < 6; ++v)
      for (int dd = 0; dd < nd; ++dd) {
        const auto& row = S.table.terms[(size_t)v * S.table.nface + dd];
        const size_t r = (size_t)v * nd + dd;
        cn[r] = (int32_t)row.size();
        for (size_t j = 0; j < row.size(); ++j) {
          ck[r * J + j] = row[j].ie | (row[j].idx << 2);
          cw[r * J + j] = row[j].w;
        }
      }
    upload(sp->ftc_n, cn, st); upload(sp->ftc_key, ck, st); upload(sp->ftc_w, cw, st);
    upload(sp->ftc_w32, std::vector<float>(cw.begin(), cw.end()), st);
    d.ndg2 = nd;
    d.ftc_J = J;
  }
  upload(sp->ghat, S.ghat, st);
  upload(sp->ghat32, std::vector<float>(S.ghat.begin(), S.ghat.end()), st);
  upload(sp->sixv, S.sixv, st);
  upload(sp->sixv32, std::vector<float>(S.sixv.begin(), S.sixv.end()), st);
  int K = S.K, Kd = S.ref.Kd, R = 3 * Kd;
  std::vector<double> dT((size_t)K * R), dN((size_t)R * K);
  for (int a = 0; a < 3; ++a)
    for (int m = 0; m < Kd; ++m)
      for (int i = 0; i < K; ++i) {
        double v = S.ref.dfac[a][(size_t)m * K + i];
        dT[(size_t)i * R + a * Kd + m] = v;
        dN[(size_t)(a * Kd + m) * K + i] = v;
      }
  std::vector<double> mT((size_t)K * K), mN(S.ref.mfac);
  for (int m = 0; m < K; ++m)
    for (int i = 0; i < K; ++i) mT[(size_t)i * K + m] = S.ref.mfac[(size_t)m * K + i];
  upload(sp->dT, dT, st); upload(sp->dN, dN, st); upload(sp->mT, mT, st); upload(sp->mN, mN, st);
  // padded factors for the MFMA kernel: Dp[a Mp + m][i], Mp = roundup(Kd, 16), Cp = roundup(K, 16)
  d.Mp = (Kd + 15) & ~15;
  d.Cp = (K + 15) & ~15;
  std::vector<double> Dp((size_t)3 * d.Mp * d.Cp, 0.0);
  for (int a = 0; a < 3; ++a)
    for (int m = 0; m < Kd; ++m)
      for (int i = 0; i < K; ++i) Dp[(size_t)(a * d.Mp + m) * d.Cp + i] = S.ref.dfac[a][(size_t)m * K + i];
  // structural zeros: d_a phi_i has degree deg(i) - 1, and psi_m is
  // orthogonal to P_d for m >= C(d+3, 3); so Dp[m][i] = 0 once
  // m >= C(max(1, deg(i) - 1) + 3, 3).  Per 16-row tile: first column of
  // GEMM 1 and row extent of GEMM 2.
  {
    auto degA = [&](int i) { auto a = S.ref.idx.mi[i]; return std::max(1, a[0] + a[1] + a[2]); };
    auto lim = [&](int i) { return ModeIndex::count_upto(std::max(1, degA(i) - 1)); };
    for (int a = 0; a < 3; ++a)
      for (int m = 0; m < Kd; ++m)
        for (int i = 0; i < K; ++i)
          if (m >= lim(i)) Dp[(size_t)(a * d.Mp + m) * d.Cp + i] = 0.0;
    // the tile tables serve the MFMA element kernels, which take K <= 192
    // (M <= 8); higher orders run the LDS element kernel and leave them unused
    int Mt = (Kd + 15) / 16, Rt = (K + 15) / 16;
    if (Mt > 16 || Rt > 16) Mt = Rt = 0;
    for (int t = 0; t < Mt; ++t) {
      int kb = K;
      for (int i = 0; i < K; ++i)
        if (lim(i) > 16 * t) { kb = i; break; }
      d.kbeg1[t] = kb & ~3;
    }
    for (int t2 = 0; t2 < Rt; ++t2) {
      int ke = 0;
      for (int i = 16 * t2; i < std::min(K, 16 * t2 + 16); ++i) ke = std::max(ke, std::min(lim(i), Kd));
      d.kend2[t2] = (ke + 3) & ~3;
    }
  }
  upload(sp->Dp64, Dp, st);
  upload(sp->Dp32, std::vector<float>(Dp.begin(), Dp.end()), st);
  {
    // reference mass matrix M = mfac^T mfac (the scalar kernel applies the
    // two factors in turn), zero padded to Cp x Cp
    std::vector<double> Mm((size_t)d.Cp * d.Cp, 0.0);
    for (int i = 0; i < K; ++i)
      for (int j = 0; j < K; ++j) {
        double s = 0.0;
        for (int m = 0; m < K; ++m) s += S.ref.mfac[(size_t)m * K + i] * S.ref.mfac[(size_t)m * K + j];
        Mm[(size_t)i * d.Cp + j] = s;
      }
    upload(sp->Mm64, Mm, st);
    upload(sp->Mm32, std::vector<float>(Mm.begin(), Mm.end()), st);
  }
  // fp16-split factors: scaled by a power of two so that |D| < 1, each
  // entry as hi + lo fp16; GEMM 1 layout [3][Mp][K32], GEMM 2 layout
  // [K16][3 Mh] (transposed, components padded to Mh rows)
  {
    d.K32 = (K + 31) & ~31;
    d.Mh = (Kd + 31) & ~31;
    d.R2g = 3 * d.Mh;
    d.R2 = d.Mh + 8;
    const int K16 = (K + 15) & ~15;
    double amax = 0;
    for (double v : Dp) amax = std::max(amax, std::fabs(v));
    int ex = 0;
    std::frexp(amax > 0 ? amax : 1.0, &ex);
    const double sD = std::ldexp(1.0, -ex);
    d.dinv = (float)std::ldexp(1.0, ex);
    auto split = [&](double v, uint16_t& hi, uint16_t& lo) {
      float s = (float)(v * sD);
      _Float16 h = (_Float16)s;
      _Float16 l = (_Float16)(s - (float)h);
      std::memcpy(&hi, &h, 2);
      std::memcpy(&lo, &l, 2);
    };
    std::vector<uint16_t> h1((size_t)3 * d.Mp * d.K32, 0), l1(h1.size(), 0);
    for (int a = 0; a < 3; ++a)
      for (int m = 0; m < d.Mp; ++m)
        for (int k = 0; k < K; ++k) {
          size_t o = ((size_t)a * d.Mp + m) * d.K32 + k;
          split(Dp[(size_t)(a * d.Mp + m) * d.Cp + k], h1[o], l1[o]);
        }
    std::vector<uint16_t> h2((size_t)K16 * d.R2g, 0), l2(h2.size(), 0);
    for (int i = 0; i < K; ++i)
      for (int a = 0; a < 3; ++a)
        for (int m = 0; m < d.Mp; ++m) {
          size_t o = (size_t)i * d.R2g + a * d.Mh + m;
          split(Dp[(size_t)(a * d.Mp + m) * d.Cp + i], h2[o], l2[o]);
        }
    upload(sp->Dh1hi, h1, st); upload(sp->Dh1lo, l1, st);
    upload(sp->Dh2hi, h2, st); upload(sp->Dh2lo, l2, st);
    for (int t = 0; t < 16; ++t) {
      d.kbeg1h[t] = d.kbeg1[t] & ~31;
      d.kend2h[t] = d.kend2[t];
    }
  }
  upload(sp->dT32, std::vector<float>(dT.begin(), dT.end()), st);
  upload(sp->dN32, std::vector<float>(dN.begin(), dN.end()), st);
  upload(sp->mT32, std::vector<float>(mT.begin(), mT.end()), st);
  upload(sp->mN32, std::vector<float>(mN.begin(), mN.end()), st);
  std::vector<int32_t> scp(S.sc_ptr.size());
  if (S.sc_ptr.back() >= (1LL << 31)) throw std::invalid_argument("scatter map too large");
  for (size_t i = 0; i < scp.size(); ++i) scp[i] = (int32_t)S.sc_ptr[i];
  d.sc_chunk_max = 0;
  for (int64_t g0 = 0; g0 < S.ndof; g0 += 256) {
    int64_t g1 = std::min<int64_t>(S.ndof, g0 + 256);
    d.sc_chunk_max = std::max(d.sc_chunk_max, (int)(S.sc_ptr[g1] - S.sc_ptr[g0]));
  }
  upload(sp->sc_ptr, scp, st); upload(sp->sc_code, S.sc_code, st); upload(sp->sc_w, S.sc_w, st);
  upload(sp->sc_w32, std::vector<float>(S.sc_w.begin(), S.sc_w.end()), st);
  sp->cmask.assign(S.ndof, 0);
  upload_mask(sp.get(), st);
  std::vector<double> xe((size_t)nt * 12);
  for (int e = 0; e < nt; ++e)
    for (int k = 0; k < 4; ++k)
      for (int c = 0; c < 3; ++c) xe[(size_t)e * 12 + 3 * k + c] = xyz[3 * (size_t)tet[4 * e + k] + c];
  upload(sp->xe, xe, st);
  upload(sp->pts3, sp->p3, st);
  d.ntile = tile_count(nt);
  d.Kp = (K + 3) & ~3;
  sp->yl64.alloc((size_t)d.ntile * TILE * K * sizeof(double));
  sp->yl32.alloc((size_t)d.ntile * TILE * K * sizeof(float));
  sp->xt64.alloc((size_t)d.ntile * TILE * d.Kp * sizeof(double));
  sp->xt32.alloc((size_t)d.ntile * TILE * d.Kp * sizeof(float));
  sp->partial.alloc((size_t)4 * (reduce_blocks((int)S.ndof) + 16) * sizeof(double));
  sp->scal.alloc(sizeof(CgScalars));
  d.desc = sp->desc.as<ElemDesc>(); d.slot = sp->slot.as<int32_t>();
  d.ftc_n = sp->ftc_n.as<int32_t>(); d.ftc_key = sp->ftc_key.as<int32_t>();
  d.ftc_w = sp->ftc_w.as<double>(); d.ftc_w32 = sp->ftc_w32.as<float>();
  d.ft_ptr = sp->ft_ptr.as<int32_t>(); d.ft_key = sp->ft_key.as<int32_t>(); d.ft_w = sp->ft_w.as<double>();
  d.ghat = sp->ghat.as<double>(); d.ghat32 = sp->ghat32.as<float>();
  d.sixv = sp->sixv.as<double>(); d.sixv32 = sp->sixv32.as<float>();
  d.dT = sp->dT.as<double>(); d.dN = sp->dN.as<double>(); d.dT32 = sp->dT32.as<float>(); d.dN32 = sp->dN32.as<float>();
  d.mT = sp->mT.as<double>(); d.mN = sp->mN.as<double>(); d.mT32 = sp->mT32.as<float>(); d.mN32 = sp->mN32.as<float>();
  d.sc_ptr = sp->sc_ptr.as<int32_t>(); d.sc_code = sp->sc_code.as<int32_t>();
  d.sc_w = sp->sc_w.as<double>(); d.sc_w32 = sp->sc_w32.as<float>();
  d.diag = sp->diag_d.as<double>(); d.invdiag = sp->invdiag.as<double>(); d.invdiag32 = sp->invdiag32.as<float>();
  d.Dp64 = sp->Dp64.as<double>(); d.Dp32 = sp->Dp32.as<float>();
  d.Mm64 = sp->Mm64.as<double>(); d.Mm32 = sp->Mm32.as<float>();
  d.xt64 = sp->xt64.as<double>(); d.xt32 = sp->xt32.as<float>();
  d.Dh1hi = sp->Dh1hi.as<uint16_t>(); d.Dh1lo = sp->Dh1lo.as<uint16_t>();
  d.Dh2hi = sp->Dh2hi.as<uint16_t>(); d.Dh2lo = sp->Dh2lo.as<uint16_t>();
  if (sp->slab.ok) {
    const tsem::Slab& L = sp->slab;
    upload(sp->sl_meta, L.meta, st); upload(sp->sl_chunk, L.chunk, st); upload(sp->sl_eoff, L.eoff, st);
    upload(sp->sl_perm, L.perm, st); upload(sp->sl_tk, L.tk, st); upload(sp->sl_tw, L.tw, st);
    upload(sp->sl_tw32, std::vector<float>(L.tw.begin(), L.tw.end()), st);
    d.sl_n1 = L.n1; d.sl_n2 = L.n2; d.sl_JT = L.JT; d.sl_nrow = L.nrow; d.sl_chunk_max = L.chunk_max;
    d.sl_diag = L.diag;
    for (int v = 0; v < 6; ++v) d.sl_sgn[v] = L.sgn[v];
    d.sl_meta = sp->sl_meta.as<int32_t>(); d.sl_chunk = sp->sl_chunk.as<int32_t>();
    d.sl_eoff = sp->sl_eoff.as<int32_t>(); d.sl_perm = sp->sl_perm.as<int32_t>();
    d.sl_tk = sp->sl_tk.as<uint16_t>(); d.sl_tw = sp->sl_tw.as<double>(); d.sl_tw32 = sp->sl_tw32.as<float>();
  }
  d.hlds_v = gemm_h_colmax_offset(d);  // after the slab fields: their tables share the region
  // assembled diagonal on the device from the scatter map, host copy for
  // the boundary-condition setup
  sp->diag_d.alloc((size_t)S.ndof * 8);
  sp->invdiag.alloc((size_t)S.ndof * 8);
  sp->invdiag32.alloc((size_t)S.ndof * 4);
  d.diag = sp->diag_d.as<double>(); d.invdiag = sp->invdiag.as<double>(); d.invdiag32 = sp->invdiag32.as<float>();
  {
    const size_t KK = (size_t)S.K * S.K;
    std::vector<double> g2(9 * KK);
    for (int ab = 0; ab < 9; ++ab) std::copy(S.ref.grad2[ab].begin(), S.ref.grad2[ab].end(), g2.begin() + ab * KK);
    upload(sp->grad2, g2, st);
  }
  launch_diagonal(d, sp->grad2.as<double>(), sp->diag_d.as<double>(), sp->invdiag.as<double>(),
                  sp->invdiag32.as<float>(), st);
  HIPCHK(hipGetLastError());
  sp->diag.resize(S.ndof);
  HIPCHK(hipMemcpyAsync(sp->diag.data(), sp->diag_d.p, (size_t)S.ndof * 8, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  ctx->nspaces++;
  *out = sp.release();
  API_END
}

int tsem_space_destroy(tsem_space* sp) {
  API_BEGIN
  if (sp) {
    tsem_ctx* c = sp->ctx;
    if (c) (void)hipStreamSynchronize(c->stream);
    delete sp;
    if (c && --c->nspaces == 0 && c->closing) ctx_release(c);
  }
  API_END
}

int tsem_space_diagonal_host(const tsem_space* sp, double* diag) {
  API_BEGIN
  if (!sp || !diag) throw std::invalid_argument("null argument");
  std::memcpy(diag, sp->diag.data(), sp->diag.size() * 8);
  API_END
}

int tsem_space_info(const tsem_space* sp, int64_t* info) {
  API_BEGIN
  if (!sp || !info) throw std::invalid_argument("null argument");
  const Space& S = sp->h;
  info[0] = S.ndof; info[1] = S.K; info[2] = S.topo.nv; info[3] = S.topo.ne; info[4] = S.topo.nf; info[5] = S.topo.nt;
  for (int k = 0; k < 4; ++k) info[6 + k] = S.ndg[k];
  std::vector<std::pair<int64_t, double>> tm;
  int64_t n = 0;
  for (int e = 0; e < S.topo.nt; ++e)
    for (int i = 0; i < S.K; ++i) { tm.clear(); S.gather_terms(e, i, tm); n += (int64_t)tm.size(); }
  info[10] = n;
  info[11] = sp->nq3;
  API_END
}

int tsem_space_topology(const tsem_space* sp, int32_t* ev, int32_t* fv, int32_t* fe, int32_t* tf) {
  API_BEGIN
  if (!sp) throw std::invalid_argument("null space");
  const Topology& T = sp->h.topo;
  if (ev) std::memcpy(ev, T.edge_vtx.data(), T.edge_vtx.size() * 4);
  if (fv) std::memcpy(fv, T.face_vtx.data(), T.face_vtx.size() * 4);
  if (fe) std::memcpy(fe, T.face_edge.data(), T.face_edge.size() * 4);
  if (tf) std::memcpy(tf, T.tet_face.data(), T.tet_face.size() * 4);
  API_END
}

int tsem_space_maps(const tsem_space* sp, int32_t* off, int32_t* ind, double* val) {
  API_BEGIN
  if (!sp || !off || !ind || !val) throw std::invalid_argument("null argument");
  const Space& S = sp->h;
  std::vector<std::pair<int64_t, double>> tm;
  int64_t n = 0;
  off[0] = 0;
  for (int e = 0; e < S.topo.nt; ++e)
    for (int i = 0; i < S.K; ++i) {
      tm.clear();
      S.gather_terms(e, i, tm);
      for (auto& pr : tm) { ind[n] = (int32_t)pr.first; val[n] = pr.second; ++n; }
      off[(size_t)e * S.K + i + 1] = (int32_t)n;
    }
  API_END
}

int tsem_space_slab(const tsem_space* sp, int64_t* out) {
  API_BEGIN
  if (!sp || !out) throw std::invalid_argument("null argument");
  const DevSpace& d = sp->d;
  out[0] = (d.sl_meta && slab_enabled() && mfma_h_supported(d) && slab_h_fits(d) && slab_scatter_fits(d) &&
            apply_path() == 0) ? 1 : 0;
  out[1] = sp->slab.ok ? sp->slab.total : 0;
  out[2] = sp->slab.chunk_max;
  out[3] = sp->slab.JT;
  API_END
}

int tsem_space_slab_apply_host(const tsem_space* sp, const double* yl, double* y) {
  API_BEGIN
  if (!sp || !yl || !y) throw std::invalid_argument("null argument");
  tsem::slab_apply_host(sp->h, sp->slab, yl, y);
  API_END
}

int tsem_space_dof_start(const tsem_space* sp, int64_t* start) {
  API_BEGIN
  if (!sp || !start) throw std::invalid_argument("null argument");
  std::memcpy(start, sp->h.dof_start.data(), sp->h.dof_start.size() * 8);
  API_END
}

int tsem_space_sum_dof(const tsem_space* sp, int32_t* sd) {
  API_BEGIN
  if (!sp || !sd) throw std::invalid_argument("null argument");
  std::memcpy(sd, sp->h.sum_dof.data(), sp->h.sum_dof.size() * 4);
  API_END
}

int tsem_space_ref_matrices(const tsem_space* sp, double* mass, double* stiff6, double* pw) {
  API_BEGIN
  if (!sp) throw std::invalid_argument("null space");
  const Space& S = sp->h;
  size_t KK = (size_t)S.K * S.K;
  if (mass) std::memcpy(mass, S.ref.mass.data(), KK * 8);
  if (stiff6) for (int k = 0; k < 6; ++k) std::memcpy(stiff6 + k * KK, S.ref.stiff[k].data(), KK * 8);
  if (pw) std::memcpy(pw, S.pairw.data(), S.pairw.size() * 8);
  API_END
}

int tsem_space_factors(const tsem_space* sp, double* dfac, double* mfac, double* grad2) {
  API_BEGIN
  if (!sp) throw std::invalid_argument("null space");
  const RefElement& R = sp->h.ref;
  size_t n = (size_t)R.Kd * R.K;
  if (dfac) for (int a = 0; a < 3; ++a) std::memcpy(dfac + a * n, R.dfac[a].data(), n * 8);
  if (mfac) std::memcpy(mfac, R.mfac.data(), (size_t)R.K * R.K * 8);
  if (grad2) for (int k = 0; k < 9; ++k) std::memcpy(grad2 + k * (size_t)R.K * R.K, R.grad2[k].data(), (size_t)R.K * R.K * 8);
  API_END
}

int tsem_set_marks(tsem_space* sp, const int32_t* mv, const int32_t* me, const int32_t* mf) {
  API_BEGIN
  if (!sp || !mv || !me || !mf) throw std::invalid_argument("null argument");
  Space& S = sp->h;
  S.mark[0].assign(mv, mv + S.topo.nv);
  S.mark[1].assign(me, me + S.topo.ne);
  S.mark[2].assign(mf, mf + S.topo.nf);
  API_END
}

// geometries of dimension dim carrying boundary mark `flag`, in index order
static std::vector<int> marked(const Space& S, int dim, int flag) {
  std::vector<int> out;
  if (S.mark[dim].empty()) throw std::logic_error("boundary marks not set");
  for (size_t g = 0; g < S.mark[dim].size(); ++g)
    if (S.mark[dim][g] == flag) out.push_back((int)g);
  return out;
}

int tsem_bc_count(const tsem_space* sp, int flag, int64_t* n3) {
  API_BEGIN
  if (!sp || !n3) throw std::invalid_argument("null argument");
  for (int d = 0; d < 3; ++d) n3[d] = (int64_t)marked(sp->h, d, flag).size();
  API_END
}

int tsem_bc_points(const tsem_space* sp, int flag, double* pv, double* pe, double* pf) {
  API_BEGIN
  if (!sp) throw std::invalid_argument("null space");
  const Space& S = sp->h;
  int q = sp->nq1, q2 = sp->nq2;
  if (pv) {
    auto G = marked(S, 0, flag);
    for (size_t j = 0; j < G.size(); ++j)
      for (int c = 0; c < 3; ++c) pv[3 * j + c] = S.xyz[3 * (size_t)G[j] + c];
  }
  if (pe) {
    auto G = marked(S, 1, flag);
    for (size_t j = 0; j < G.size(); ++j) {
      const double* a = &S.xyz[3 * (size_t)S.edge_node[G[j]][0]];
      const double* b = &S.xyz[3 * (size_t)S.edge_node[G[j]][1]];
      for (int k = 0; k < q; ++k)
        for (int c = 0; c < 3; ++c) pe[(j * q + k) * 3 + c] = a[c] * (1 - sp->t1[k]) + b[c] * sp->t1[k];
    }
  }
  if (pf) {
    auto G = marked(S, 2, flag);
    for (size_t j = 0; j < G.size(); ++j) {
      const double* n0 = &S.xyz[3 * (size_t)S.face_node[G[j]][0]];
      const double* n1 = &S.xyz[3 * (size_t)S.face_node[G[j]][1]];
      const double* n2 = &S.xyz[3 * (size_t)S.face_node[G[j]][2]];
      for (int k = 0; k < q2; ++k) {
        const double* bq = &sp->p2[3 * (size_t)k];
        for (int c = 0; c < 3; ++c) pf[(j * q2 + k) * 3 + c] = n0[c] * bq[2] + n1[c] * bq[0] + n2[c] * bq[1];
      }
    }
  }
  API_END
}

int tsem_qp_coords(tsem_space* sp, int e0, int e1, double* out) {
  API_BEGIN
  if (!sp || !out || e0 < 0 || e1 > sp->h.topo.nt || e0 > e1) throw std::invalid_argument("bad element range");
  launch_qp_coords(sp->d, sp->xe.as<double>(), e0, e1, sp->nq3, sp->pts3.as<double>(), out, sp->st());
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(sp->st()));
  API_END
}

int tsem_rhs_begin(tsem_space* sp) {
  API_BEGIN
  if (!sp) throw std::invalid_argument("null space");
  const Space& S = sp->h;
  if (!sp->wphi.p) {
    // w_q phi_i(x_q) depends on the order and the rule only: built once per
    // (order, qorder) in the process
    static std::mutex mu;
    static std::map<std::pair<int, int>, std::shared_ptr<const std::vector<double>>> cache;
    std::shared_ptr<const std::vector<double>> tab;
    {
      std::lock_guard<std::mutex> lk(mu);
      auto it = cache.find({S.M, sp->qorder});
      if (it != cache.end()) tab = it->second;
    }
    if (!tab) {
      auto w = std::make_shared<std::vector<double>>((size_t)sp->nq3 * S.K);
      std::vector<long double> val(S.K);
      for (int q = 0; q < sp->nq3; ++q) {
        S.ref.eval(sp->p3[3 * q], sp->p3[3 * q + 1], sp->p3[3 * q + 2], val.data(), nullptr);
        for (int i = 0; i < S.K; ++i) (*w)[(size_t)q * S.K + i] = sp->w3[q] * (double)val[i];
      }
      std::lock_guard<std::mutex> lk(mu);
      tab = cache.emplace(std::make_pair(S.M, sp->qorder), w).first->second;
    }
    upload(sp->wphi, *tab, sp->st());
    HIPCHK(hipStreamSynchronize(sp->st()));  // the table is read from host memory
  }
  size_t nbl = (size_t)sp->d.ntile * TILE * S.K;
  sp->bl.alloc(nbl * sizeof(double));
  HIPCHK(hipMemsetAsync(sp->bl.p, 0, nbl * sizeof(double), sp->st()));
  HIPCHK(hipStreamSynchronize(sp->st()));
  sp->rhs_open = true;
  API_END
}

int tsem_rhs_add(tsem_space* sp, int e0, int e1, const double* fq) {
  API_BEGIN
  if (!sp || !fq) throw std::invalid_argument("null argument");
  if (!sp->rhs_open) { g_err = "tsem_rhs_begin not called"; return TSEM_ESTATE; }
  if (e0 < 0 || e1 > sp->h.topo.nt || e0 > e1) throw std::invalid_argument("bad element range");
  const int ns = rhs_splits(sp->d, e1 - e0, sp->nq3);
  if (ns > 1) sp->rhs_part.alloc((size_t)ns * (e1 - e0) * sp->h.K * sizeof(double));
  launch_rhs_local(sp->d, e0, e1, sp->nq3, sp->wphi.as<double>(), fq, sp->bl.as<double>(),
                   ns > 1 ? sp->rhs_part.as<double>() : nullptr, sp->st());
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(sp->st()));
  API_END
}

int tsem_rhs_end(tsem_space* sp, double* b) {
  API_BEGIN
  if (!sp || !b) throw std::invalid_argument("null argument");
  if (!sp->rhs_open) { g_err = "tsem_rhs_begin not called"; return TSEM_ESTATE; }
  launch_scatter<double>(sp->d, sp->bl.as<double>(), b, sp->st());
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(sp->st()));
  sp->rhs_open = false;
  API_END
}

// ---------------------------------------------------------------- BCs ------
// tsem_neumann and tsem_dirichlet are in bc.hip

static void apply_dev(tsem_space* sp, double stiff, double mass, int constrained, const double* x, double* y,
                      double* partial) {
  OpCoef op;
  op.stiff = stiff; op.mass = mass; op.constrained = constrained;
  launch_apply<double>(sp->d, op, x, y, sp->yl64.as<double>(), partial, sp->st());
  HIPCHK(hipGetLastError());
}

int tsem_zero_bc(tsem_space* sp, int flag) {
  API_BEGIN
  if (!sp) throw std::invalid_argument("null space");
  Space& S = sp->h;
  if (S.mark[0].empty()) { g_err = "boundary marks not set"; return TSEM_ESTATE; }
  bool grown = false;  // DOFs eliminated now that were free before
  for (int dim = 0; dim <= 2; ++dim) {
    int ng = dim == 0 ? S.topo.nv : dim == 1 ? S.topo.ne : S.topo.nf;
    for (int g = 0; g < ng; ++g)
      if (S.mark[dim][g] == flag)
        for (int i = 0; i < S.ndg[dim]; ++i) {
          uint8_t& m = sp->cmask[S.geo_dof(dim, g) + i];
          grown |= !m;
          m = 1;
        }
  }
  upload_mask(sp, sp->st());
  HIPCHK(hipStreamSynchronize(sp->st()));
  if (grown) sp->drop_pmg();  // the levels and graphs depend on the mask only
  API_END
}

int tsem_zero_rows(tsem_space* sp, int flag, double* x_dev) {
  API_BEGIN
  if (!sp || !x_dev) throw std::invalid_argument("null argument");
  Space& S = sp->h;
  if (S.mark[0].empty()) { g_err = "boundary marks not set"; return TSEM_ESTATE; }
  std::vector<uint8_t> m(S.ndof, 0);
  for (int dim = 0; dim <= 2; ++dim) {
    int ng = dim == 0 ? S.topo.nv : dim == 1 ? S.topo.ne : S.topo.nf;
    for (int g = 0; g < ng; ++g)
      if (S.mark[dim][g] == flag)
        for (int i = 0; i < S.ndg[dim]; ++i) m[S.geo_dof(dim, g) + i] = 1;
  }
  DevBuf md;
  upload(md, m, sp->st());
  launch_block_mask((int)S.ndof, 1, md.as<uint8_t>(), nullptr, x_dev, sp->st());
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(sp->st()));
  API_END
}

int tsem_apply(tsem_space* sp, double stiff, double mass, int constrained, int prec, const void* x, void* y) {
  API_BEGIN
  if (!sp || !x || !y) throw std::invalid_argument("null argument");
  OpCoef op;
  op.stiff = stiff; op.mass = mass; op.constrained = constrained;
  if (prec == 64) launch_apply<double>(sp->d, op, (const double*)x, (double*)y, sp->yl64.as<double>(), nullptr, sp->st());
  else if (prec == 32) launch_apply<float>(sp->d, op, (const float*)x, (float*)y, sp->yl32.as<float>(), nullptr, sp->st());
  else throw std::invalid_argument("prec must be 64 or 32");
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(sp->st()));
  API_END
}

int tsem_gather(tsem_space* sp, const double* x, double* xl) {
  API_BEGIN
  if (!sp || !x || !xl) throw std::invalid_argument("null argument");
  launch_gather(sp->d, x, xl, 0, sp->st());
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(sp->st()));
  API_END
}

int tsem_scatter(tsem_space* sp, const double* yl, double* y) {
  API_BEGIN
  if (!sp || !yl || !y) throw std::invalid_argument("null argument");
  launch_scatter<double>(sp->d, yl, y, sp->st());
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(sp->st()));
  API_END
}

int tsem_timing(int enable) {
  API_BEGIN
  timing_enable(enable != 0);
  API_END
}

int tsem_timing_read(double* out) {
  API_BEGIN
  if (!out) throw std::invalid_argument("null argument");
  timing_read(out);
  API_END
}

}  // extern "C"

bool tsem::graphs_enabled() { return g_pcg_graphs; }

static bool g_block_vcycle = [] {
  const char* v = getenv("TSEM_BLOCK_VCYCLE");
  return !(v && v[0] == '0');
}();
bool tsem::block_vcycle_enabled() { return g_block_vcycle; }

extern "C" {

int tsem_set_block_vcycle(int enable) {
  API_BEGIN
  g_block_vcycle = enable != 0;
  API_END
}

int tsem_set_graphs(int enable) {
  API_BEGIN
  g_pcg_graphs = enable != 0;
  API_END
}

int tsem_set_slab(int enable) {
  API_BEGIN
  set_slab(enable != 0);
  API_END
}

int tsem_set_apply_path(int path) {
  API_BEGIN
  if (path < 0 || path > 2) throw std::invalid_argument("path must be 0 (auto), 1 (scalar) or 2 (no fp16 split)");
  set_apply_path(path);
  API_END
}

int tsem_diagonal(tsem_space* sp, double* dd) {
  API_BEGIN
  if (!sp || !dd) throw std::invalid_argument("null argument");
  HIPCHK(hipMemcpyAsync(dd, sp->diag_d.p, sp->h.ndof * 8, hipMemcpyDeviceToDevice, sp->st()));
  HIPCHK(hipStreamSynchronize(sp->st()));
  API_END
}

}  // extern "C"

// ------------------------------------------------------------- solvers ------
namespace {

// Preconditioned CG, reference formulation (src/TSEMSolver.cpp:1775-1930):
// r = A x - b, y = P r, p = -y, alpha = r.y / p.Ap, x += alpha p,
// r += alpha Ap, beta = r.y_new / r.y, p = beta p - y.
// stop: criterion 0 -> |r|_1 < tol, 1 -> |r|_2 / bnorm < tol (recursive r).
// Check hook: called with the scalars read back after every chunk of
// iterations and a `refresh(restart)` that recomputes z = P r and r.z for the
// current r (for a hook that replaced r) and, with restart, also resets the
// search direction to p = -z; returns the residual measure tested against tol.
using CheckHook = std::function<double(CgScalars&, const std::function<void(bool)>&)>;

template <typename T>
int pcg_core(tsem_space* sp, const T* b, T* x, double tol, int criterion, double bnorm, int max_iter, int precond,
             int smooth, int check_every, int* iters, double* l1_out, double* l2_out,
             const CheckHook* hook = nullptr) {
  const DevSpace& d = sp->d;
  int n = d.ndof;
  hipStream_t st = sp->st();
  DevBuf* V = (sizeof(T) == 8) ? sp->v64 : sp->v32;
  for (int k = 0; k < 4; ++k) V[k].alloc((size_t)n * sizeof(T));
  T* r = V[0].as<T>(); T* z = V[1].as<T>(); T* p = V[2].as<T>(); T* q = V[3].as<T>();
  T* yl = (sizeof(T) == 8) ? (T*)sp->yl64.p : (T*)sp->yl32.p;
  CgScalars* sc = sp->scal.as<CgScalars>();
  double* part = sp->partial.as<double>();
  int nb = reduce_blocks(n);
  OpCoef op;
  op.constrained = 1;
  // r = A x - b
  launch_apply<T>(d, op, x, r, yl, nullptr, st);
  launch_axpby<T>(n, T(-1), b, T(1), r, st);
  // initial preconditioned residual and scalars: the update kernel with
  // alpha = 0 computes z = D^-1 r (Jacobi) and the partial sums
  CgScalars zero{};
  HIPCHK(hipMemcpyAsync(sc, &zero, sizeof(CgScalars), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemsetAsync(p, 0, (size_t)n * sizeof(T), st));
  HIPCHK(hipMemsetAsync(q, 0, (size_t)n * sizeof(T), st));
  // x += alpha p, r += alpha q, z = P r; partials [r.z | |r|_1 | |r|_2^2] then
  // beta = r.z / r.z_old in one final reduction.  With pq_part, alpha is
  // formed from the p.q partials inside the update; else it is read from sc
  // finish = false leaves beta to k_p_update_beta
  auto update_and_precondition = [&](const double* pq_part, bool finish) {
    launch_cg_update<T>(d, n, x, r, precond == 4 ? nullptr : z, p, q, sc, part, precond == 4 ? 0 : 1, pq_part,
                        st);
    if (precond == 4) {
      sp->pmg->vcycle<T>(r, z, smooth, st);
      launch_dot<T>(n, r, z, part, st);  // replaces the r.z partials
    }
    if (finish) launch_final_beta(part, nb, 1, sc, st);
  };
  update_and_precondition(nullptr, true);
  // p = -z
  launch_axpby<T>(n, T(-1), z, T(0), p, st);
  CgScalars hs{};
  HIPCHK(hipMemcpyAsync(&hs, sc, sizeof(CgScalars), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  auto resid = [&](const CgScalars& s) {
    return criterion == 0 ? s.l1 : std::sqrt(s.l2) / bnorm;
  };
  double res = resid(hs);
  int it = 0;
  static const bool trace = getenv("TSEM_PCG_TRACE") != nullptr;
  if (trace) fprintf(stderr, "[tsem pcg] start res %.3e bnorm %.3e rz %.3e\n", res, bnorm, hs.rz_new);
  // the p.q partials go to the fourth quarter of the partial buffer, which
  // the update's own partials do not reach
  double* part_pq = part + 3 * (size_t)nb;
  auto iteration = [&]() {
    if (g_fused_alpha) {
      launch_apply<T>(d, op, p, q, yl, part_pq, st);
      update_and_precondition(part_pq, !g_fused_beta);
    } else {
      launch_apply<T>(d, op, p, q, yl, part, st);
      launch_final_alpha(part, nb, sc, st);
      update_and_precondition(nullptr, true);
    }
    if (g_fused_alpha && g_fused_beta) launch_p_update_beta<T>(n, p, z, sc, part, st);
    else launch_p_update<T>(n, p, z, sc, st);
  };
  auto& GS = (sizeof(T) == 8) ? sp->graph64 : sp->graph32;
  double last_chunk = 0.0;  // host seconds of the previous chunk
  // chunk length: check_every iterations, shortened near the end to the
  // iterations the observed rate of the last chunk predicts (rounded down to
  // a power of two, so few graphs are captured), which saves most of the
  // overshoot past the tolerance
  const int full = std::max(1, check_every);
  int nrun = full;
  double res_prev = res;
  int run_prev = 0;
  while (res >= tol && it <= max_iter) {
    background_wait(sp->ctx, std::max(last_chunk, 1e-4));
    const auto chunk_t0 = std::chrono::steady_clock::now();
    nrun = full;
    if (run_prev > 0 && res < res_prev && res > 0) {
      const double rate = std::log(res / res_prev) / run_prev;  // < 0
      const double left = std::log(tol / res) / rate;           // iterations still needed
      if (left < full) {
        int q = 1;
        while (2 * q <= (int)left) q *= 2;
        nrun = std::max(1, std::min(full, q));
      }
    }
    int gslot = 0;  // slot 0: the full chunk; slot j + 1: a chunk of 2^j iterations
    if (nrun != full) {
      int j = 0;
      while ((1 << j) < nrun) ++j;
      gslot = j + 1;
      if (gslot >= tsem_space::GRAPH_SLOTS) {
        nrun = full;
        gslot = 0;
      }
    }
    auto& G = GS.slot[gslot];
    std::vector<const void*> key;
    if (g_pcg_graphs && !timing_on() && it + nrun <= max_iter + 1) {
      // the chunk of nrun iterations as one captured graph, replayed while
      // the pointers and parameters it captured stay the same
      key = {x, r, z, p, q, yl, part, sc, (const void*)(uintptr_t)sp->state_gen, (const void*)(intptr_t)precond,
             (const void*)(intptr_t)smooth, (const void*)(intptr_t)nrun, (const void*)(intptr_t)apply_path(),
             (const void*)(intptr_t)slab_enabled()};
      if ((!G.exec || G.key != key) && G.failed != key) {
        // a capture can be invalidated from outside (another thread's
        // device-wide synchronization): its launches never ran, so the chunk
        // then runs eagerly and this key is not captured again
        G.reset();
        hipGraph_t g = nullptr;
        bool ok = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) == hipSuccess;
        if (ok) {
          try {
            for (int k = 0; k < nrun; ++k) iteration();
          } catch (...) {
            ok = false;
          }
          if (hipStreamEndCapture(st, &g) != hipSuccess) ok = false;
          if (ok && g && hipGraphInstantiate(&G.exec, g, nullptr, nullptr, 0) != hipSuccess) {
            G.exec = nullptr;
            ok = false;
          }
          if (g) (void)hipGraphDestroy(g);
        }
        if (ok && G.exec) {
          G.key = key;
        } else {
          G.reset();
          G.failed = key;
          (void)hipGetLastError();
        }
      }
    }
    if (!key.empty() && G.exec && G.key == key) {
      HIPCHK(hipGraphLaunch(G.exec, st));
      it += nrun;
    } else {
      for (int k = 0; k < nrun && it <= max_iter; ++k) {
        iteration();
        ++it;
      }
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(&hs, sc, sizeof(CgScalars), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    last_chunk = std::chrono::duration<double>(std::chrono::steady_clock::now() - chunk_t0).count();
    res_prev = res;
    run_prev = nrun;
    if (hook) {
      auto refresh = [&](bool restart) {
        static const double zero = 0.0;
        HIPCHK(hipMemcpyAsync(&sc->alpha, &zero, sizeof(double), hipMemcpyHostToDevice, st));
        update_and_precondition(nullptr, true);
        if (restart) launch_axpby<T>(n, T(-1), z, T(0), p, st);
        HIPCHK(hipMemcpyAsync(&hs, sc, sizeof(CgScalars), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
      };
      res = (*hook)(hs, refresh);
    } else {
      res = resid(hs);
    }
    if (trace)
      fprintf(stderr, "[tsem pcg%s] it %d res %.3e l1 %.3e l2 %.3e rz %.3e pq %.3e alpha %.3e beta %.3e\n",
              sizeof(T) == 8 ? "64" : "32", it, res, hs.l1, std::sqrt(hs.l2), hs.rz_new, hs.pq, hs.alpha, hs.beta);
    if (!std::isfinite(res)) break;
  }
  *iters = it;
  *l1_out = hs.l1;
  *l2_out = std::sqrt(hs.l2);
  return (res < tol) ? TSEM_OK : TSEM_ENOCONV;
}

double true_relres(tsem_space* sp, const double* b, const double* x, double* r, double* bn) {
  int n = sp->d.ndof;
  hipStream_t st = sp->st();
  DevBuf& t = sp->v64[5];
  t.alloc((size_t)n * 8);
  apply_dev(sp, 1.0, 0.0, 1, x, t.as<double>(), nullptr);
  double* part = sp->partial.as<double>();
  int nb = reduce_blocks(n);
  launch_residual(sp->d, n, b, t.as<double>(), r, part, st);
  CgScalars* sc = sp->scal.as<CgScalars>();
  launch_reduce_final(part, nb, 2, sc->tmp + 4, st);
  double h[2];
  HIPCHK(hipMemcpyAsync(h, sc->tmp + 4, 16, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (bn) *bn = std::sqrt(h[1]);
  return std::sqrt(h[0]) / std::sqrt(h[1]);
}

}  // namespace

extern "C" {

int tsem_pcg(tsem_space* sp, const double* b, double* x, const tsem_pcg_opts* o, tsem_pcg_stats* stats) {
  API_BEGIN
  if (!sp || !b || !x || !o) throw std::invalid_argument("null argument");
  if (o->precond != 1 && o->precond != 4) throw std::invalid_argument("precond must be 1 (Jacobi) or 4 (p-MG)");
  if (o->mixed < 0 || o->mixed > 2) throw std::invalid_argument("mixed must be 0, 1 or 2");
  ForegroundScope fg(sp->ctx);
  if (!(o->tol > 0)) throw std::invalid_argument("tol must be positive");
  int n = sp->d.ndof;
  hipStream_t st = sp->st();
  if (o->precond == 4 && !sp->pmg) sp->pmg.reset(new Pmg(sp, st));
  auto t0 = std::chrono::steady_clock::now();
  tsem_pcg_stats s{};
  int rc = TSEM_OK;
  sp->v64[4].alloc((size_t)n * 8);
  double* r64 = sp->v64[4].as<double>();
  double bnorm = 0;
  if (!o->mixed) {
    double rr = true_relres(sp, b, x, r64, &bnorm);
    (void)rr;
    if (bnorm == 0) bnorm = 1;
    int it = 0;
    double l1 = 0, l2 = 0;
    rc = pcg_core<double>(sp, b, x, o->tol, o->criterion, bnorm, o->max_iter, o->precond, o->smooth_steps,
                          o->check_every, &it, &l1, &l2);
    s.iterations = it;
    s.outer = 1;
    s.l1 = l1;
    s.relres = true_relres(sp, b, x, r64, nullptr);
  } else if (o->mixed == 1) {
    // FP32 PCG with reliable updates: one FP32 Krylov recursion for the whole
    // solve (vectors scaled by 1 / ||b||); whenever its residual has fallen by
    // the factor inner_tol since the last update, and when it reports
    // convergence, the FP32 correction is added to the FP64 x, the true
    // residual b - A x is formed in FP64 and replaces the FP32 one.  The
    // search direction is kept while the replaced residual differs from the
    // recursive one by at most half its norm (measured, |r_rec - r_true|);
    // a larger gap
    // (the FP32 recursion has drifted) restarts it from p = -z, since the
    // stale direction belongs to a different residual.  A breakdown of the
    // FP32 recursion (non-finite residual) returns to the last FP64 state with
    // a restarted direction and updates every 10^-2 of reduction (at most 3
    // times)
    for (int k = 0; k < 4; ++k) sp->v32[k].alloc((size_t)n * 4);
    sp->v32[4].alloc((size_t)n * 4);
    sp->v32[5].alloc((size_t)n * 4);
    float* rhs32 = sp->v32[4].as<float>();
    float* d32 = sp->v32[5].as<float>();
    float* r32 = sp->v32[0].as<float>();  // the residual pcg_core<float> iterates on
    double rel = true_relres(sp, b, x, r64, &bnorm);
    if (bnorm == 0) bnorm = 1;
    double delta = std::min(0.5, std::max(o->inner_tol, 1e-6));
    int updates = 0, restarts = 0, recoveries = 0;
    float* p32 = sp->v32[2].as<float>();
    float* q32 = sp->v32[3].as<float>();
    CgScalars* scd = sp->scal.as<CgScalars>();
    double* part = sp->partial.as<double>();
    const int nbr = reduce_blocks(n);
    static const bool trace = getenv("TSEM_PCG_TRACE") != nullptr;
    if (rel > o->tol) {
      launch_cast_f64_f32(n, r64, 1.0 / bnorm, rhs32, st);
      launch_zero<float>(n, d32, st);
      double rmax = rel;
      bool failed = false;
      CheckHook hook = [&](CgScalars& hs, const std::function<void(bool)>& refresh) -> double {
        const double rr = std::sqrt(hs.l2);  // recursive residual relative to ||b||
        if (!std::isfinite(rr)) {
          if (++recoveries > 3) { failed = true; return 0.0; }
          launch_zero<float>(n, d32, st);
          launch_zero<float>(n, p32, st);
          launch_zero<float>(n, q32, st);
          launch_cast_f64_f32(n, r64, -1.0 / bnorm, r32, st);
          refresh(true);
          ++restarts;
          delta = std::max(delta, 1e-2);
          rmax = rel;
          if (trace) fprintf(stderr, "[tsem mixed] FP32 breakdown: back to true residual %.3e\n", rel);
          return rel;
        }
        if (rr >= delta * rmax && rr >= o->tol) {
          rmax = std::max(rmax, rr);
          return rr;
        }
        launch_add_f32_f64(n, d32, bnorm, x, st);
        launch_zero<float>(n, d32, st);
        rel = true_relres(sp, b, x, r64, nullptr);
        ++updates;
        if (rel <= o->tol || !std::isfinite(rel) || updates > o->max_iter) {
          failed = !(rel <= o->tol);
          return rel <= o->tol ? rel : 0.0;  // 0 ends the loop
        }
        // the recursion's residual is A d - rhs = -(b - A x) / ||b||; the gap
        // |r_rec - r_true| before it is replaced
        launch_gap(n, r32, r64, 1.0 / bnorm, part, st);
        launch_reduce_final(part, nbr, 1, &scd->tmp[6], st);
        double g2 = 0;
        HIPCHK(hipMemcpyAsync(&g2, &scd->tmp[6], sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        const double gap = std::sqrt(g2);
        launch_cast_f64_f32(n, r64, -1.0 / bnorm, r32, st);
        const bool restart = !(gap <= 0.5 * rel);
        refresh(restart);
        restarts += restart ? 1 : 0;
        if (trace)
          fprintf(stderr, "[tsem mixed] update %d: recursive %.3e true %.3e gap %.3e restarts %d\n", updates, rr,
                  rel, gap, restarts);
        rmax = rel;
        return rel;
      };
      int it = 0;
      double l1 = 0, l2 = 0;
      pcg_core<float>(sp, rhs32, d32, o->tol, 1, 1.0, o->inner_max, o->precond, o->smooth_steps, o->check_every,
                      &it, &l1, &l2, &hook);
      // an iteration limit leaves a correction behind
      launch_add_f32_f64(n, d32, bnorm, x, st);
      rel = true_relres(sp, b, x, r64, nullptr);
      s.iterations = it;
      if (failed || !(rel <= o->tol)) rc = TSEM_ENOCONV;
    }
    s.outer = updates;
    s.relres = rel;
    s.l1 = 0;
  } else {
    // FP64 iterative refinement around restarted FP32 inner PCG
    for (int k = 0; k < 4; ++k) sp->v32[k].alloc((size_t)n * 4);
    sp->v32[4].alloc((size_t)n * 4);
    sp->v32[5].alloc((size_t)n * 4);
    float* r32 = sp->v32[4].as<float>();
    float* d32 = sp->v32[5].as<float>();
    double rel = true_relres(sp, b, x, r64, &bnorm);
    if (bnorm == 0) bnorm = 1;
    int outer = 0, total = 0;
    rc = TSEM_OK;
    while (rel > o->tol) {
      if (outer >= o->max_iter) { rc = TSEM_ENOCONV; break; }
      double rn = rel * bnorm;
      launch_cast_f64_f32(n, r64, 1.0 / rn, r32, st);
      HIPCHK(hipMemsetAsync(d32, 0, (size_t)n * 4, st));
      int it = 0;
      double l1 = 0, l2 = 0;
      // the inner solve needs no more than the reduction still missing (with
      // a margin of 3 for the FP32 operator); never less than inner_tol
      double itol = std::min(0.5, std::max(o->inner_tol, 0.3 * o->tol / rel));
      pcg_core<float>(sp, r32, d32, itol, 1, 1.0, o->inner_max, o->precond, o->smooth_steps,
                      o->check_every, &it, &l1, &l2);
      total += it;
      launch_add_f32_f64(n, d32, rn, x, st);
      rel = true_relres(sp, b, x, r64, nullptr);
      ++outer;
      if (!std::isfinite(rel)) { rc = TSEM_ENOCONV; break; }
    }
    s.iterations = total;
    s.outer = outer;
    s.relres = rel;
    s.l1 = 0;
  }
  HIPCHK(hipStreamSynchronize(st));
  s.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (stats) *stats = s;
  if (rc != TSEM_OK) { g_err = "PCG did not reach the tolerance"; return rc; }
  API_END
}

}  // extern "C"
