// Served stream of Poisson, eigen and system requests behind the C ABI
// (tsem_serve_*).
//
// A server owns `slots` worker threads.  Each worker takes the smallest
// submitted request of its lane by estimated work (small-request workers take
// only small requests, the others the smallest large one, else the smallest
// small one)
// as soon as the request's device footprint fits the budget, and solves it
// on a space of its own context (stream), so requests in flight overlap on
// the device.  Spaces are cached per caller topology key: a later request
// with the same key only replaces the vertex coordinates
// (tsem_space_set_geometry) and rebuilds its load vector and boundary data;
// idle cached spaces are dropped, least recently used first, when the budget
// needs their memory.  With classes on, contexts of small requests are
// foreground and those of large ones background (tsem_ctx_set_class), so a
// large solve pauses between its chunks of iterations while a small one runs.
//
// A Poisson request is the mixed-BC problem of
// examples/poisson_mixBC/poisson.cpp:74-89: load vector, Neumann load on mark
// 2, Dirichlet elimination on mark 1, PCG to a relative residual.  Its data
// come from the device evaluator of the manufactured family (tsem_manufactured,
// at quadrature and boundary points on the request's stream) or through
// callbacks: f at device points on the request's stream, u and du/dn at host
// points.  An eigen request takes the same mesh path, then zero Dirichlet
// data on its Dirichlet parts and LOBPCG; a system request is a PCG solve on
// a space the caller owns (no pool entry, no footprint).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <random>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "handle.hpp"

namespace {

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

int64_t binom3(int M) { return (int64_t)(M + 1) * (M + 2) * (M + 3) / 6; }

// device bytes of one space with its request buffers, estimated from the
// element count and order: DOFs from the Kuhn-mesh geometry ratios (faces
// 2 nt, edges 7/6 nt, vertices nt / 6), scatter map, element-local buffers,
// slab tables, solver vectors and quadrature points
int64_t footprint(int nt, int M, int q) {
  const double n1 = M - 1, n2 = (M - 1) * (M - 2) / 2.0, n3 = (M - 1) * (M - 2) * (M - 3) / 6.0;
  const double N = nt * (n3 + 2 * n2 + 7.0 / 6.0 * n1 + 1.0 / 6.0);
  const double K = (double)binom3(M);
  const double E = nt;
  const double pts = std::min(2e6, E * q * q * q) + E * q * q * q / 8.0;
  return (int64_t)(3.2 * N * 16 + E * K * 32 + E * 300 + N * 170 + pts * 32) + (64LL << 20);
}

struct Job {
  int64_t id = 0;
  // the request, with copies of its arrays
  tsem_serve_request req{};
  std::vector<double> xyz;
  std::vector<int32_t> tet, labels;
  double t_submit = 0;
  tsem_manufactured_def mdef{};
  // filled in by the worker
  bool done = false;
  tsem_serve_result res{};
  std::string error;
  std::vector<double> x;
};

struct Entry {
  tsem_ctx* ctx = nullptr;
  tsem_space* sp = nullptr;
  bool busy = false;
  double last = 0;
  int64_t bytes = 0;
  // per-space request buffers (device), and the outward normals of the
  // current request's Neumann faces (host, per face)
  double *b = nullptr, *x = nullptr, *pts = nullptr, *fq = nullptr;
  int64_t n_pts = 0;
  std::vector<double> normals;
  // boundary points, normals and values on the device (manufactured data)
  double* bcb = nullptr;
  int64_t bcb_n = 0;
  double* bc_dev(int64_t n) {
    if (n > bcb_n) {
      if (bcb) (void)hipFree(bcb);
      bcb = nullptr;
      bcb_n = 0;
      if (hipMalloc(&bcb, (size_t)n * 8) != hipSuccess) throw std::bad_alloc();
      bcb_n = n;
    }
    return bcb;
  }
  void release() {
    if (b) (void)hipFree(b);
    if (x) (void)hipFree(x);
    if (pts) (void)hipFree(pts);
    if (fq) (void)hipFree(fq);
    if (bcb) (void)hipFree(bcb);
    b = x = pts = fq = bcb = nullptr;
    bcb_n = 0;
    if (sp) tsem_space_destroy(sp);
    if (ctx) tsem_ctx_destroy(ctx);
    sp = nullptr;
    ctx = nullptr;
  }
};

// pool key: caller topology, order, quadrature, sizes (a key of 0 is never
// shared: the request id takes the place of the element count)
typedef std::tuple<int64_t, int, int, int, int64_t> Key;

// a failed library call inside a request, with its status
struct CallError : std::runtime_error {
  int rc;
  CallError(int r, const std::string& m) : std::runtime_error(m), rc(r) {}
};
void check(int rc, const char* what) {
  if (rc != TSEM_OK) throw CallError(rc, std::string(what) + ": " + tsem_last_error());
}

}  // namespace

struct tsem_server {
  int device = 0;
  tsem_serve_opts o{};
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Job*> pending;
  std::map<int64_t, std::unique_ptr<Job>> jobs;
  std::map<Key, std::vector<std::unique_ptr<Entry>>> pool;
  int64_t in_use = 0, next_id = 1;
  bool closing = false;
  std::vector<std::thread> workers;

  double cost(const Job& j) const {
    if (j.req.kind == TSEM_REQ_SYSTEM) return (double)j.req.space->h.topo.nt * binom3(j.req.space->h.M);
    return (double)j.req.nt * binom3(j.req.M);
  }
  bool small(const Job& j) const { return cost(j) <= o.small_cost; }
  Key key(const Job& j) const {
    return Key(j.req.topology, j.req.M, j.req.qorder, j.req.nv, j.req.topology ? (int64_t)j.req.nt : -j.id);
  }

  // with the lock held: a busy pool entry for the job (an idle cached space,
  // or a new entry within the budget after evicting idle spaces), or null
  Entry* reserve(const Job& j) {
    const Key k = key(j);
    if (j.req.topology)
      for (auto& e : pool[k])
        if (!e->busy) {
          e->busy = true;
          return e.get();
        }
    int64_t need = footprint(j.req.nt, j.req.M, j.req.qorder);
    if (j.req.kind == TSEM_REQ_EIGEN) {
      // LOBPCG's blocks: [X P W] and their images, residuals, FP32 V-cycle columns
      const double N = (double)j.req.nt * j.req.M * j.req.M * j.req.M / 6.0 + j.req.nv;
      need += (int64_t)(N * j.req.nev * (11 * 8 + 40));
    }
    while (in_use + need > o.hbm_budget) {
      Entry* victim = nullptr;
      Key vk;
      for (auto& kv : pool)
        for (auto& e : kv.second)
          if (!e->busy && (!victim || e->last < victim->last)) { victim = e.get(); vk = kv.first; }
      if (!victim) {
        bool any_busy = false;
        for (auto& kv : pool)
          for (auto& e : kv.second) any_busy |= e->busy;
        if (any_busy) return nullptr;
        break;  // an idle device takes any request
      }
      in_use -= victim->bytes;
      victim->release();
      auto& lst = pool[vk];
      lst.erase(std::find_if(lst.begin(), lst.end(), [&](const std::unique_ptr<Entry>& p) { return p.get() == victim; }));
    }
    auto e = std::make_unique<Entry>();
    e->busy = true;
    e->bytes = need;
    in_use += need;
    Entry* r = e.get();
    pool[k].push_back(std::move(e));
    return r;
  }

  // next job of worker w's lane and its pool entry; false once closing and drained
  bool next(int w, Job*& job, Entry*& ent) {
    std::unique_lock<std::mutex> lk(mu);
    const bool small_lane = w < o.small_slots;
    for (;;) {
      // small lane: the smallest small request; general lane: the smallest
      // large one, else the smallest small one (shortest job first by the
      // work estimate; submission order among equals)
      Job* pick = nullptr;
      auto smallest = [&](bool want_small) {
        Job* best = nullptr;
        for (Job* j : pending)
          if (small(*j) == want_small && (!best || cost(*j) < cost(*best))) best = j;
        return best;
      };
      if (!small_lane) pick = smallest(false);
      if (!pick) pick = smallest(true);
      if (pick) {
        // a system request runs on the caller's space: no pool entry
        Entry* e = pick->req.kind == TSEM_REQ_SYSTEM ? nullptr : reserve(*pick);
        if (e || pick->req.kind == TSEM_REQ_SYSTEM) {
          pending.erase(std::find(pending.begin(), pending.end(), pick));
          job = pick;
          ent = e;
          return true;
        }
      }
      if (closing && pending.empty()) return false;
      cv.wait_for(lk, std::chrono::milliseconds(50));
    }
  }

  void solve(Job& j, Entry& e);
  void prepare(Job& j, Entry& e);
  void poisson(Job& j, Entry& e, bool fg);
  void eigen(Job& j, Entry& e);
  void system(Job& j);
  void run(int w);
};

// outward unit normals of the faces with mark 2, in tsem_bc_points order
static void neumann_normals(const Space& S, std::vector<double>& out) {
  out.clear();
  std::vector<int> owner(S.topo.nf, -1);
  for (int t = 0; t < S.topo.nt; ++t)
    for (int k = 0; k < 4; ++k) owner[S.topo.tet_face[4 * (size_t)t + k]] = 4 * t + k;
  for (int f = 0; f < S.topo.nf; ++f) {
    if (S.mark[2][f] != 2) continue;
    const double* a = &S.xyz[3 * (size_t)S.face_node[f][0]];
    const double* b = &S.xyz[3 * (size_t)S.face_node[f][1]];
    const double* c = &S.xyz[3 * (size_t)S.face_node[f][2]];
    const int t = owner[f] / 4, k = owner[f] % 4;
    const double* o = &S.xyz[3 * (size_t)S.tet[4 * (size_t)t + k]];  // the vertex opposite face k
    double u[3], v[3], n[3];
    for (int d = 0; d < 3; ++d) { u[d] = b[d] - a[d]; v[d] = c[d] - a[d]; }
    n[0] = u[1] * v[2] - u[2] * v[1];
    n[1] = u[2] * v[0] - u[0] * v[2];
    n[2] = u[0] * v[1] - u[1] * v[0];
    const double len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    const double side = n[0] * (o[0] - a[0]) + n[1] * (o[1] - a[1]) + n[2] * (o[2] - a[2]);
    const double s = (side > 0 ? -1.0 : 1.0) / len;
    for (int d = 0; d < 3; ++d) out.push_back(n[d] * s);
  }
}

// the request's space: built, or a cached one with this request's
// coordinates; then this request's boundary marks
void tsem_server::prepare(Job& j, Entry& e) {
  const tsem_serve_request& r = j.req;
  const bool fg = o.classes && small(j);
  if (!e.sp) {
    j.res.built = 1;
    check(tsem_ctx_create(device, &e.ctx), "context");
    if (o.classes) check(tsem_ctx_set_class(e.ctx, fg ? 1 : 2), "solve class");
    check(tsem_space_create(e.ctx, r.M, r.nv, j.xyz.data(), r.nt, j.tet.data(), nullptr, r.qorder, &e.sp), "space");
    const Space& S = e.sp->h;
    const int64_t N = S.ndof;
    e.n_pts = std::max<int64_t>(1, std::min<int64_t>((int64_t)r.nt * e.sp->nq3, 2000000));
    e.n_pts = std::max<int64_t>(e.n_pts, e.sp->nq3);
    if (hipMalloc(&e.b, N * 8) != hipSuccess || hipMalloc(&e.x, N * 8) != hipSuccess ||
        hipMalloc(&e.pts, e.n_pts * 24) != hipSuccess || hipMalloc(&e.fq, e.n_pts * 8) != hipSuccess)
      throw std::bad_alloc();
  } else {
    check(tsem_space_set_geometry(e.sp, j.xyz.data()), "geometry");
  }
  // boundary marks of this request (a cached space may have served other
  // labels or Dirichlet / Neumann parts): a geometry lies on a labelled
  // part when all its vertices do; Dirichlet (1) wins over Neumann (2)
  const Space& S = e.sp->h;
  auto mark = [&](int32_t bits) { return (bits & r.dirichlet) ? 1 : (bits & r.neumann) ? 2 : 0; };
  std::vector<int32_t> mv(S.topo.nv), me(S.topo.ne), mf(S.topo.nf);
  auto lab = [&](int v) { return j.labels.empty() ? 0 : j.labels[v]; };
  for (int v = 0; v < S.topo.nv; ++v) mv[v] = mark(lab(v));
  for (int g = 0; g < S.topo.ne; ++g)
    me[g] = mark(lab(S.topo.edge_vtx[2 * (size_t)g]) & lab(S.topo.edge_vtx[2 * (size_t)g + 1]));
  for (int g = 0; g < S.topo.nf; ++g)
    mf[g] = mark(lab(S.topo.face_vtx[3 * (size_t)g]) & lab(S.topo.face_vtx[3 * (size_t)g + 1]) &
                 lab(S.topo.face_vtx[3 * (size_t)g + 2]));
  check(tsem_set_marks(e.sp, mv.data(), me.data(), mf.data()), "marks");
}

void tsem_server::solve(Job& j, Entry& e) {
  const double t0 = now_s();
  j.res.queued_s = t0 - j.t_submit;
  prepare(j, e);
  j.res.space_s = now_s() - t0;
  const bool fg = o.classes && small(j);
  if (fg) check(tsem_foreground_begin(e.ctx), "foreground");
  try {
    if (j.req.kind == TSEM_REQ_EIGEN)
      eigen(j, e);
    else
      poisson(j, e, fg);
  } catch (...) {
    if (fg) (void)tsem_foreground_end(e.ctx);
    throw;
  }
  if (fg) check(tsem_foreground_end(e.ctx), "foreground");
}

void tsem_server::poisson(Job& j, Entry& e, bool fg) {
  const tsem_serve_request& r = j.req;
  tsem_space* sp = e.sp;
  tsem_ctx* ctx = e.ctx;
  hipStream_t st = ctx->stream;
  const Space& S = sp->h;
  const bool man = r.manufactured != nullptr;
  // values of the manufactured problem at host points (and normals): up to
  // the device, evaluated there, back
  auto man_host = [&](int what, const std::vector<double>& pts, const std::vector<double>* nrm, double* out) {
    const int64_t n = (int64_t)pts.size() / 3;
    if (n == 0) return;
    double* d = e.bc_dev(n * (nrm ? 7 : 4));
    HIPCHK(hipMemcpyAsync(d, pts.data(), (size_t)n * 24, hipMemcpyHostToDevice, st));
    if (nrm) HIPCHK(hipMemcpyAsync(d + 3 * n, nrm->data(), (size_t)n * 24, hipMemcpyHostToDevice, st));
    double* v = d + (nrm ? 6 : 3) * n;
    check(tsem_manufactured(&j.mdef, what, n, d, nrm ? d + 3 * n : nullptr, v, (void*)st), "manufactured values");
    HIPCHK(hipMemcpyAsync(out, v, (size_t)n * 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
  };
  const double t1 = now_s();
  // load vector in pieces of at most n_pts points; background requests
  // pause between them while foreground work runs
  check(tsem_background_wait(ctx, 0.05), "wait");
  check(tsem_rhs_begin(sp), "load vector");
  const int per = std::max(1, (int)(e.n_pts / sp->nq3));
  for (int e0 = 0; e0 < r.nt; e0 += per) {
    const int e1 = std::min(r.nt, e0 + per);
    const int64_t n = (int64_t)(e1 - e0) * sp->nq3;
    check(tsem_qp_coords(sp, e0, e1, e.pts), "quadrature points");
    if (man)
      check(tsem_manufactured(&j.mdef, 1, n, e.pts, nullptr, e.fq, (void*)st), "manufactured load");
    else if (r.f(r.user, e.pts, n, e.fq, (void*)st) != 0)
      throw std::invalid_argument("load callback failed");
    check(tsem_rhs_add(sp, e0, e1, e.fq), "load vector");
    if (!fg) check(tsem_background_wait(ctx, 0.01), "wait");
  }
  check(tsem_rhs_end(sp, e.b), "load vector");
  int64_t c2[3], c1[3];
  check(tsem_bc_count(sp, 2, c2), "boundary");
  if (c2[2] > 0) {
    if (!man && !r.g) throw std::invalid_argument("Neumann faces but no flux callback");
    const int64_t n = c2[2] * sp->nq2;
    std::vector<double> pf(3 * n), nrm(3 * n), gq(n);
    check(tsem_bc_points(sp, 2, nullptr, nullptr, pf.data()), "boundary points");
    neumann_normals(S, e.normals);  // the coordinates may have changed
    for (int64_t f = 0; f < c2[2]; ++f)
      for (int k = 0; k < sp->nq2; ++k)
        for (int d = 0; d < 3; ++d) nrm[3 * (f * sp->nq2 + k) + d] = e.normals[3 * f + d];
    if (man)
      man_host(2, pf, &nrm, gq.data());
    else if (r.g(r.user, pf.data(), nrm.data(), n, gq.data()) != 0)
      throw std::invalid_argument("flux callback failed");
    check(tsem_neumann(sp, gq.data(), 2, e.b), "Neumann load");
  }
  check(tsem_bc_count(sp, 1, c1), "boundary");
  if (c1[0] + c1[1] + c1[2] > 0) {
    if (!man && !r.u) throw std::invalid_argument("Dirichlet geometries but no value callback");
    std::vector<double> pv(3 * c1[0]), pe(3 * c1[1] * sp->nq1), pf(3 * c1[2] * sp->nq2);
    std::vector<double> uv(c1[0]), ue(c1[1] * sp->nq1), uf(c1[2] * sp->nq2);
    check(tsem_bc_points(sp, 1, pv.data(), pe.data(), pf.data()), "boundary points");
    if (man) {
      man_host(0, pv, nullptr, uv.data());
      man_host(0, pe, nullptr, ue.data());
      man_host(0, pf, nullptr, uf.data());
    } else if (r.u(r.user, pv.data(), (int64_t)uv.size(), uv.data()) != 0 ||
               r.u(r.user, pe.data(), (int64_t)ue.size(), ue.data()) != 0 ||
               r.u(r.user, pf.data(), (int64_t)uf.size(), uf.data()) != 0) {
      throw std::invalid_argument("value callback failed");
    }
    check(tsem_dirichlet(sp, uv.data(), ue.data(), uf.data(), 1, e.b, nullptr), "Dirichlet");
  }
  const double t2 = now_s();
  check(tsem_background_wait(ctx, 0.05), "wait");
  HIPCHK(hipMemsetAsync(e.x, 0, S.ndof * 8, st));
  tsem_pcg_opts po{};
  po.tol = r.tol > 0 ? r.tol : 1e-10;
  po.criterion = 1;
  po.max_iter = 100000;
  po.precond = r.precond ? r.precond : 1;
  po.mixed = r.mixed;
  po.inner_tol = 1e-2;
  po.inner_max = 100000;
  po.smooth_steps = 2;
  po.check_every = 10;
  tsem_pcg_stats ps{};
  const int rc = tsem_pcg(sp, e.b, e.x, &po, &ps);
  if (rc != TSEM_OK && rc != TSEM_ENOCONV) check(rc, "PCG");
  j.x.resize(S.ndof);
  HIPCHK(hipMemcpyAsync(j.x.data(), e.x, S.ndof * 8, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  j.res.status = rc;
  j.res.n_dof = S.ndof;
  j.res.iterations = ps.iterations;
  j.res.outer = ps.outer;
  j.res.relres = ps.relres;
  j.res.setup_s = t2 - t1;
  j.res.solve_s = now_s() - t2;
}

void tsem_server::eigen(Job& j, Entry& e) {
  const tsem_serve_request& r = j.req;
  tsem_space* sp = e.sp;
  hipStream_t st = e.ctx->stream;
  const int64_t N = sp->h.ndof;
  const int m = r.nev;
  const double t1 = now_s();
  check(tsem_zero_bc(sp, 1), "zero Dirichlet data");
  std::vector<double> X0((size_t)N * m);
  std::mt19937_64 g(r.seed);
  std::normal_distribution<double> nd(0.0, 1.0);
  for (double& v : X0) v = nd(g);
  double* X = nullptr;
  if (hipMalloc(&X, X0.size() * 8) != hipSuccess) throw std::bad_alloc();
  std::vector<double> lam(m, 0.0);
  int rc = TSEM_OK;
  tsem_lobpcg_stats ls{};
  try {
    HIPCHK(hipMemcpyAsync(X, X0.data(), X0.size() * 8, hipMemcpyHostToDevice, st));
    const double t2 = now_s();
    j.res.setup_s = t2 - t1;
    tsem_lobpcg_opts lo{};
    lo.tol = r.tol > 0 ? r.tol : 1e-8;
    lo.criterion = 1;
    lo.max_iter = 2000;
    lo.precond = 4;
    lo.smooth_steps = 4;
    lo.stiff = 1.0;
    rc = tsem_lobpcg(sp, m, X, lam.data(), &lo, &ls);
    if (rc != TSEM_OK && rc != TSEM_ENOCONV) check(rc, "LOBPCG");
    j.x.resize(X0.size());
    HIPCHK(hipMemcpyAsync(j.x.data(), X, X0.size() * 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    j.res.solve_s = now_s() - t2;
  } catch (...) {
    (void)hipFree(X);
    throw;
  }
  (void)hipFree(X);
  if (r.eigenvalues) std::copy(lam.begin(), lam.end(), r.eigenvalues);
  j.res.status = rc;
  j.res.n_dof = N;
  j.res.iterations = ls.iterations;
  j.res.outer = ls.applies;
  j.res.relres = ls.max_resid;
}

void tsem_server::system(Job& j) {
  const tsem_serve_request& r = j.req;
  const double t0 = now_s();
  j.res.queued_s = t0 - j.t_submit;
  tsem_pcg_opts po{};
  po.tol = r.tol > 0 ? r.tol : 1e-10;
  po.criterion = 1;
  po.max_iter = 100000;
  po.precond = r.precond ? r.precond : 1;
  po.mixed = r.mixed;
  po.inner_tol = 1e-2;
  po.inner_max = 100000;
  po.smooth_steps = 2;
  po.check_every = 10;
  tsem_pcg_stats ps{};
  const int rc = tsem_pcg(r.space, r.b_dev, r.x_dev, &po, &ps);
  if (rc != TSEM_OK && rc != TSEM_ENOCONV) check(rc, "PCG");
  HIPCHK(hipStreamSynchronize(r.space->st()));
  j.res.status = rc;
  j.res.n_dof = r.space->h.ndof;
  j.res.iterations = ps.iterations;
  j.res.outer = ps.outer;
  j.res.relres = ps.relres;
  j.res.solve_s = now_s() - t0;
}

void tsem_server::run(int w) {
  (void)hipSetDevice(device);
  Job* j = nullptr;
  Entry* e = nullptr;
  while (next(w, j, e)) {
    int status = TSEM_OK;
    std::string err;
    try {
      if (j->req.kind == TSEM_REQ_SYSTEM)
        system(*j);
      else
        solve(*j, *e);
    } catch (const CallError& ex) {
      status = ex.rc;
      err = ex.what();
    } catch (const std::bad_alloc&) {
      status = TSEM_ENOMEM;
      err = "allocation failed";
    } catch (const std::invalid_argument& ex) {
      status = TSEM_EINVAL;
      err = ex.what();
    } catch (const std::exception& ex) {
      status = TSEM_EDEVICE;
      err = ex.what();
    }
    std::lock_guard<std::mutex> lk(mu);
    if (status != TSEM_OK) {
      j->res.status = status;
      j->error = err;
    }
    if (!e) {
      // a system request: the space is the caller's
    } else if (status != TSEM_OK || !j->req.topology) {
      // a failed request's space may be half set up, and a space without a
      // topology key is never reused: free it now
      in_use -= e->bytes;
      e->release();
      auto& lst = pool[key(*j)];
      lst.erase(std::find_if(lst.begin(), lst.end(), [&](const std::unique_ptr<Entry>& p) { return p.get() == e; }));
    } else {
      e->busy = false;
      e->last = now_s();
    }
    j->res.total_s = now_s() - j->t_submit;
    j->done = true;
    cv.notify_all();
  }
}

int tsem_serve_open(int device, const tsem_serve_opts* opts, tsem_server** out) {
  API_BEGIN
  if (!out) throw std::invalid_argument("null output");
  int n = 0;
  HIPCHK(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) throw std::invalid_argument("device index out of range");
  auto s = std::make_unique<tsem_server>();
  s->device = device;
  if (opts) {
    s->o = *opts;
  } else {
    s->o.small_slots = -1;
    s->o.classes = 1;
  }
  // 4 workers by default: at a fixed arrival rate, fewer requests in flight
  // interfere less (2,000-request stream at 6.9 requests/s: latency over
  // isolated p99 13.8 with 4 workers, 20.6 with 8, 29-43 with 16-24), for
  // about 10% less saturated throughput
  if (s->o.slots <= 0) s->o.slots = 4;
  if (s->o.slots > 64) throw std::invalid_argument("at most 64 workers");
  if (s->o.small_slots < 0) s->o.small_slots = s->o.slots / 2;
  if (s->o.small_slots > s->o.slots) throw std::invalid_argument("small_slots > slots");
  if (s->o.small_cost <= 0) s->o.small_cost = 4e6;
  if (s->o.hbm_budget <= 0) {
    HIPCHK(hipSetDevice(device));
    size_t fr = 0, tot = 0;
    HIPCHK(hipMemGetInfo(&fr, &tot));
    s->o.hbm_budget = (int64_t)(0.7 * (double)fr);
  }
  tsem_server* raw = s.get();
  for (int w = 0; w < s->o.slots; ++w) s->workers.emplace_back([raw, w] { raw->run(w); });
  *out = s.release();
  API_END
}

int tsem_serve_submit(tsem_server* s, const tsem_serve_request* r, int64_t* id) {
  API_BEGIN
  if (!s || !r || !id) throw std::invalid_argument("null argument");
  if (r->kind != TSEM_REQ_POISSON && r->kind != TSEM_REQ_EIGEN && r->kind != TSEM_REQ_SYSTEM)
    throw std::invalid_argument("unknown request kind");
  if (r->mixed < 0 || r->mixed > 2) throw std::invalid_argument("mixed must be 0, 1 or 2");
  if (r->precond != 0 && r->precond != 1 && r->precond != 4) throw std::invalid_argument("precond must be 1 or 4");
  auto j = std::make_unique<Job>();
  j->req = *r;
  if (r->kind == TSEM_REQ_SYSTEM) {
    if (!r->space || !r->b_dev || !r->x_dev) throw std::invalid_argument("system request needs space, b and x");
    if (!r->space->ctx) throw std::invalid_argument("system request on a host-only space");
    j->req.xyz = nullptr;
    j->req.tet = nullptr;
    j->req.labels = nullptr;
    j->req.manufactured = nullptr;
    j->t_submit = now_s();
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->closing) throw std::logic_error("server is closing");
    j->id = s->next_id++;
    *id = j->id;
    s->pending.push_back(j.get());
    s->jobs[j->id] = std::move(j);
    s->cv.notify_all();
    return TSEM_OK;
  }
  if (r->M < 1 || r->M > 12) throw std::invalid_argument("order must be 1..12");
  if (r->nv < 4 || r->nt < 1 || !r->xyz || !r->tet) throw std::invalid_argument("empty mesh");
  if (r->kind == TSEM_REQ_POISSON && !r->f && !r->manufactured)
    throw std::invalid_argument("no load callback or manufactured problem");
  if (r->kind == TSEM_REQ_EIGEN && (r->nev < 1 || r->nev > 32)) throw std::invalid_argument("nev must be 1..32");
  if (r->manufactured) {
    if (r->manufactured->nterms < 1 || r->manufactured->nterms > 16)
      throw std::invalid_argument("manufactured problem needs 1..16 terms");
    j->mdef = *r->manufactured;
    j->req.manufactured = &j->mdef;
  }
  if (j->req.qorder <= 0) j->req.qorder = r->M + 3;
  j->xyz.assign(r->xyz, r->xyz + 3 * (size_t)r->nv);
  j->tet.assign(r->tet, r->tet + 4 * (size_t)r->nt);
  for (int32_t v : j->tet)
    if (v < 0 || v >= r->nv) throw std::invalid_argument("tet vertex index out of range");
  if (r->labels) j->labels.assign(r->labels, r->labels + r->nv);
  j->req.xyz = nullptr;
  j->req.tet = nullptr;
  j->req.labels = nullptr;
  j->t_submit = now_s();
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->closing) throw std::logic_error("server is closing");
  j->id = s->next_id++;
  *id = j->id;
  s->pending.push_back(j.get());
  s->jobs[j->id] = std::move(j);
  s->cv.notify_all();
  API_END
}

int tsem_serve_poll(tsem_server* s, int64_t id, int wait, double* x, tsem_serve_result* res, int* done) {
  API_BEGIN
  if (!s || !done) throw std::invalid_argument("null argument");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->jobs.find(id);
  if (it == s->jobs.end()) throw std::invalid_argument("unknown request id");
  Job* j = it->second.get();
  if (wait)
    s->cv.wait(lk, [&] { return j->done; });
  *done = j->done ? 1 : 0;
  if (!j->done) return TSEM_OK;
  if (res) *res = j->res;
  const int status = j->res.status;
  const std::string err = j->error;
  if (x) {  // the solution is taken: release the request
    if (!j->x.empty()) std::memcpy(x, j->x.data(), j->x.size() * 8);
    s->jobs.erase(it);
  }
  if (status != TSEM_OK && status != TSEM_ENOCONV) {
    g_err = err;
    return status;
  }
  API_END
}

int tsem_serve_close(tsem_server* s) {
  API_BEGIN
  if (!s) return TSEM_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->closing = true;
    s->cv.notify_all();
  }
  for (auto& t : s->workers) t.join();
  for (auto& kv : s->pool)
    for (auto& e : kv.second) e->release();
  delete s;
  API_END
}
