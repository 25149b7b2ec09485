// C++ interface of libtsem_amd.so with the class and method names of
// TSEMLib's TetrahedralSEM.h / TSEMSolver.h, over the C ABI of tsem.h.
//
//   TSEMLib                                          here
//   TSEM<double>(M, [qorder,] mesh, fem_space)       TSEM<double>(M, [qorder,] mesh)
//   tsem.build_flag_bm(mesh)                         same (marks read from the .mesh file)
//   tsem.calc_rhs(rhs, fem_space, &f)                tsem.calc_rhs(rhs, &f)
//   tsem.impose_boundary_condition(rhs, mesh, &g, 2) same (Neumann)
//   tsem.impose_boundary_condition(A, rhs, mesh, &u, true, 1)   same (Dirichlet elimination)
//   tsem.impose_zero_boundary_condition(A, 1)        same
//   tsem.calc_l2_error(x, fem_space, &u)             tsem.calc_l2_error(x, &u)
//   SparseMatrix<double>::vmult(dst, src)            same; the matrix is a matrix-free device operator
//   TSEMSolver<double>::solve_PCG(A, x, b, tol, max_step, flag, ite_step, var)       same
//   TSEMSolver<double>::solve_InversePower(x, A, B, flag, tol, max_step)             same
//   TSEMSolver<double>::solve_LOBPCG(tsem, x, A, B, flag, tol, max_step)             same (all pairs)
//
// References: include/TetrahedralSEM.h:193-273, include/TSEMSolver.h:58-98.
// Vectors live in device memory; host element access goes through
// Vector::host() / Vector::assign() copies.  Every Vector operation and every
// library call is complete when it returns (the library runs on a
// non-blocking stream of its context, so nothing else orders them).  Errors throw std::runtime_error
// with tsem_last_error() (the reference prints and returns).
#ifndef TSEM_AMD_HPP
#define TSEM_AMD_HPP

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <map>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include "tsem.h"

namespace tsem {

inline void check(int rc) {
  if (rc != TSEM_OK) throw std::runtime_error(std::string("tsem: ") + tsem_last_error());
}
inline void hip_check(hipError_t e) {
  if (e != hipSuccess) throw std::runtime_error(std::string("HIP: ") + hipGetErrorString(e));
}

// ---------------------------------------------------------------- mesh --
// AFEPack .mesh text (examples/*/cube.mesh): point count and coordinates,
// then for dimension 0..3 a geometry count and per geometry
// "index  n vertex...  n boundary...  mark".  The vertices of edges, faces
// and tetrahedra are dim-0 geometry indices, each naming one point.
struct RegularMesh {
  std::vector<double> xyz;
  std::vector<int32_t> tet;                       // 4 point ids per tetrahedron
  std::vector<std::array<int32_t, 2>> edge_vtx;   // point ids, file order
  std::vector<std::array<int32_t, 3>> face_vtx;   // point ids, file order
  std::vector<int32_t> mark_v, mark_e, mark_f;    // marks per point, per file edge, per file face

  void readMesh(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot open mesh file " + path);
    size_t nv;
    in >> nv;
    xyz.resize(3 * nv);
    for (auto& v : xyz) in >> v;
    mark_v.assign(nv, 0);
    std::vector<int32_t> point;                   // dim-0 geometry -> point id
    auto pt = [&](int32_t g) {
      if (g < 0 || (size_t)g >= point.size()) throw std::runtime_error("bad vertex index in " + path);
      return point[g];
    };
    for (int dim = 0; dim < 4; ++dim) {
      size_t n;
      in >> n;
      for (size_t g = 0; g < n && in; ++g) {
        long idx, nvtx, nbnd, mark;
        in >> idx >> nvtx;
        if (nvtx != dim + 1) throw std::runtime_error("unexpected geometry in " + path);
        std::vector<int32_t> vt(nvtx);
        for (auto& v : vt) in >> v;
        in >> nbnd;
        for (long b = 0; b < nbnd; ++b) { long t; in >> t; }
        in >> mark;
        if (dim == 0) {
          if (vt[0] < 0 || (size_t)vt[0] >= nv) throw std::runtime_error("bad point index in " + path);
          point.push_back(vt[0]);
          mark_v[vt[0]] = (int32_t)mark;
        }
        if (dim == 1) { edge_vtx.push_back({pt(vt[0]), pt(vt[1])}); mark_e.push_back((int32_t)mark); }
        if (dim == 2) { face_vtx.push_back({pt(vt[0]), pt(vt[1]), pt(vt[2])}); mark_f.push_back((int32_t)mark); }
        if (dim == 3) for (auto v : vt) tet.push_back(pt(v));
      }
    }
    if (!in) throw std::runtime_error("malformed mesh file " + path);
  }
  int n_geometry(int dim) const {
    return dim == 0 ? (int)(xyz.size() / 3) : dim == 1 ? (int)edge_vtx.size() : dim == 2 ? (int)face_vtx.size()
                                                                                          : (int)(tet.size() / 4);
  }
};

// the device a TSEM lives on: one context (stream) per process by default
class Context {
 public:
  explicit Context(int device = 0) { check(tsem_ctx_create(device, &h_)); }
  ~Context() { tsem_ctx_destroy(h_); }
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;
  tsem_ctx* get() const { return h_; }
  static Context& instance() {
    static Context c(0);
    return c;
  }

 private:
  tsem_ctx* h_ = nullptr;
};

// -------------------------------------------------------------- vector --
template <typename T>
class Vector {
 public:
  Vector() = default;
  explicit Vector(size_t n) { reinit(n); }
  Vector(const Vector& o) { *this = o; }
  Vector& operator=(const Vector& o) {
    if (this != &o) {
      reinit(o.n_);
      if (n_) copy_device(p_, o.p_, n_);
    }
    return *this;
  }
  ~Vector() { if (p_) (void)hipFree(p_); }
  void reinit(size_t n) {
    if (n != n_) {
      if (p_) (void)hipFree(p_);
      p_ = nullptr;
      n_ = n;
      if (n) hip_check(hipMalloc(&p_, n * sizeof(T)));
    }
    if (n) {
      hip_check(hipMemset(p_, 0, n * sizeof(T)));
      hip_check(hipDeviceSynchronize());
    }
  }
  // device-to-device copy, complete on return: the library works on its own
  // non-blocking stream, which the null stream does not order against
  static void copy_device(T* dst, const T* src, size_t n) {
    hip_check(hipMemcpy(dst, src, n * sizeof(T), hipMemcpyDeviceToDevice));
    hip_check(hipDeviceSynchronize());
  }
  size_t size() const { return n_; }
  T* data() { return p_; }
  const T* data() const { return p_; }
  std::vector<T> host() const {
    std::vector<T> h(n_);
    if (n_) hip_check(hipMemcpy(h.data(), p_, n_ * sizeof(T), hipMemcpyDeviceToHost));
    return h;
  }
  void assign(const std::vector<T>& h) {
    if (h.size() != n_) reinit(h.size());
    if (n_) hip_check(hipMemcpy(p_, h.data(), n_ * sizeof(T), hipMemcpyHostToDevice));
  }

  // the deal.II Vector subset the reference's callers use
  // (src/TSEMSolver.cpp:1848-1897, examples/*); host round trips, not for
  // inner loops
  T operator()(size_t i) const {
    T v;
    hip_check(hipMemcpy(&v, p_ + i, sizeof(T), hipMemcpyDeviceToHost));
    return v;
  }
  void set(size_t i, T v) { hip_check(hipMemcpy(p_ + i, &v, sizeof(T), hipMemcpyHostToDevice)); }
  T l1_norm() const {
    T s = T(0);
    for (T v : host()) s += std::fabs(v);
    return s;
  }
  T l2_norm() const {
    T s = T(0);
    for (T v : host()) s += v * v;
    return std::sqrt(s);
  }
  T operator*(const Vector& o) const {  // scalar product
    std::vector<T> a = host(), b = o.host();
    T s = T(0);
    for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
    return s;
  }
  Vector& operator*=(T a) {
    std::vector<T> v = host();
    for (T& x : v) x *= a;
    assign(v);
    return *this;
  }
  void add(T a, const Vector& o) {  // this += a o
    std::vector<T> v = host(), w = o.host();
    for (size_t i = 0; i < v.size(); ++i) v[i] += a * w[i];
    assign(v);
  }

 private:
  T* p_ = nullptr;
  size_t n_ = 0;
};

// ------------------------------------------------------- sparse matrix --
// The assembled matrices of the reference become operator handles:
// vmult applies stiff S + mass M (with the Dirichlet elimination once it is
// imposed) matrix-free on the device.
struct SparsityPattern {
  tsem_space* sp = nullptr;
};

template <typename T>
class SparseMatrix {
 public:
  SparseMatrix() = default;
  SparseMatrix(tsem_space* sp, double stiff, double mass) : sp_(sp), stiff_(stiff), mass_(mass) {}
  // SparseMatrix(tsem.stiff_matrix.get_sparsity_pattern()) then copy_from(...)
  explicit SparseMatrix(const SparsityPattern& sp) : sp_(sp.sp), stiff_(0.0), mass_(0.0) {}
  SparsityPattern get_sparsity_pattern() const { return SparsityPattern{sp_}; }
  void vmult(Vector<T>& dst, const Vector<T>& src) const {
    if (dst.size() != src.size()) dst.reinit(src.size());
    check(tsem_apply(sp_, stiff_, mass_, constrained_ ? 1 : 0, 64, src.data(), dst.data()));
  }
  size_t m() const {
    int64_t info[12];
    check(tsem_space_info(sp_, info));
    return (size_t)info[0];
  }
  void copy_from(const SparseMatrix& o) { *this = o; }
  size_t n() const { return m(); }
  // diagonal entry of a stiffness operator (the eliminated form keeps the
  // diagonal); the library holds the assembled stiffness diagonal only
  T diag_element(size_t i) const {
    if (mass_ != 0.0) throw std::invalid_argument("diag_element: stiffness operators only");
    Vector<T> d(m());
    check(tsem_diagonal(sp_, d.data()));
    return T(stiff_) * d(i);
  }
  tsem_space* space() const { return sp_; }
  double stiff() const { return stiff_; }
  double mass() const { return mass_; }
  bool constrained() const { return constrained_; }
  void set_constrained(bool c) { constrained_ = c; }

 private:
  tsem_space* sp_ = nullptr;
  double stiff_ = 1.0, mass_ = 0.0;
  bool constrained_ = false;
};

// ---------------------------------------------------------------- TSEM --
template <typename T>
class TSEM {
 public:
  TSEM(int M, RegularMesh& mesh, Context& ctx = Context::instance()) { init(M, 30, mesh, ctx); }
  TSEM(int M, int qinfo_order, RegularMesh& mesh, Context& ctx = Context::instance()) {
    init(M, qinfo_order, mesh, ctx);
  }
  ~TSEM() { tsem_space_destroy(sp_); }
  TSEM(const TSEM&) = delete;
  TSEM& operator=(const TSEM&) = delete;

  int n_dof_total = 0;
  SparseMatrix<T> stiff_matrix, mass_matrix;
  tsem_space* space() const { return sp_; }

  // TSEM::build_flag_bm(mesh): the file's boundary marks on this space's
  // geometries (matched by their vertices)
  void build_flag_bm(const RegularMesh& mesh) {
    int64_t info[12];
    check(tsem_space_info(sp_, info));
    const int nv = (int)info[2], ne = (int)info[3], nf = (int)info[4], nt = (int)info[5];
    std::vector<int32_t> ev(2 * (size_t)ne), fv(3 * (size_t)nf), fe(3 * (size_t)nf), tf(4 * (size_t)nt);
    check(tsem_space_topology(sp_, ev.data(), fv.data(), fe.data(), tf.data()));
    std::map<std::array<int32_t, 2>, int32_t> me;
    std::map<std::array<int32_t, 3>, int32_t> mf;
    for (size_t i = 0; i < mesh.edge_vtx.size(); ++i) {
      auto k = mesh.edge_vtx[i];
      std::sort(k.begin(), k.end());
      me[k] = mesh.mark_e[i];
    }
    for (size_t i = 0; i < mesh.face_vtx.size(); ++i) {
      auto k = mesh.face_vtx[i];
      std::sort(k.begin(), k.end());
      mf[k] = mesh.mark_f[i];
    }
    std::vector<int32_t> mark_v(mesh.mark_v.begin(), mesh.mark_v.end()), mark_e(ne, 0), mark_f(nf, 0);
    mark_v.resize(nv, 0);
    for (int e = 0; e < ne; ++e) {
      std::array<int32_t, 2> k = {ev[2 * e], ev[2 * e + 1]};
      std::sort(k.begin(), k.end());
      auto it = me.find(k);
      if (it != me.end()) mark_e[e] = it->second;
    }
    for (int f = 0; f < nf; ++f) {
      std::array<int32_t, 3> k = {fv[3 * f], fv[3 * f + 1], fv[3 * f + 2]};
      std::sort(k.begin(), k.end());
      auto it = mf.find(k);
      if (it != mf.end()) mark_f[f] = it->second;
    }
    check(tsem_set_marks(sp_, mark_v.data(), mark_e.data(), mark_f.data()));
    face_vtx_ = std::move(fv);
    mark_f_ = std::move(mark_f);
  }

  // TSEM::calc_rhs(rhs, fem_space, f): f evaluated on the host at the
  // quadrature points, element chunk by element chunk
  void calc_rhs(Vector<T>& rhs, double (*f)(double*)) {
    rhs.reinit(n_dof_total);
    check(tsem_rhs_begin(sp_));
    for_chunks([&](int e0, int e1, const std::vector<double>& pts, std::vector<double>& val) {
      for (size_t q = 0; q < val.size(); ++q) val[q] = f(const_cast<double*>(&pts[3 * q]));
      Vector<double> d(val.size());
      d.assign(val);
      check(tsem_rhs_add(sp_, e0, e1, d.data()));
    });
    check(tsem_rhs_end(sp_, rhs.data()));
  }

  // Neumann data g (the normal derivative, a function of the point) on the
  // faces marked flag: impose_boundary_condition(rhs, mesh, g, flag)
  void impose_boundary_condition(Vector<T>& rhs, const RegularMesh&, double (*g)(double*), int flag = 2) {
    std::vector<double> pv, pe, pf;
    bc_points(flag, pv, pe, pf);
    std::vector<double> gq(pf.size() / 3);
    for (size_t q = 0; q < gq.size(); ++q) gq[q] = g(&pf[3 * q]);
    check(tsem_neumann(sp_, gq.data(), flag, rhs.data()));
  }

  // Dirichlet data u on the geometries marked flag: the modal interpolation
  // enters rhs and A switches to the eliminated form
  // (impose_boundary_condition(A, rhs, mesh, u, true, flag))
  void impose_boundary_condition(SparseMatrix<T>& A, Vector<T>& rhs, const RegularMesh&, double (*u)(double*),
                                 bool = true, int flag = 1) {
    std::vector<double> pv, pe, pf;
    bc_points(flag, pv, pe, pf);
    auto ev = [&](std::vector<double>& p) {
      std::vector<double> v(p.size() / 3);
      for (size_t q = 0; q < v.size(); ++q) v[q] = u(&p[3 * q]);
      return v;
    };
    std::vector<double> uv = ev(pv), ue = ev(pe), uf = ev(pf);
    check(tsem_dirichlet(sp_, uv.data(), ue.data(), uf.data(), flag, rhs.data(), nullptr));
    A.set_constrained(true);
  }

  // impose_zero_boundary_condition(A, flag)
  void impose_zero_boundary_condition(SparseMatrix<T>& A, int flag = 1) {
    check(tsem_zero_bc(sp_, flag));
    A.set_constrained(true);
  }
  // impose_zero_boundary_condition(rhs, flag): zero the DOFs of those geometries
  void impose_zero_boundary_condition(Vector<T>& x, int flag = 1) { check(tsem_zero_rows(sp_, flag, x.data())); }

  // calc_l2_error(x, fem_space, u)
  double calc_l2_error(const Vector<T>& x, double (*u)(double*)) {
    double tot = 0.0;
    for_chunks([&](int e0, int e1, const std::vector<double>& pts, std::vector<double>& val) {
      for (size_t q = 0; q < val.size(); ++q) val[q] = u(const_cast<double*>(&pts[3 * q]));
      Vector<double> d(val.size());
      d.assign(val);
      double s = 0;
      check(tsem_l2_error_chunk(sp_, x.data(), e0, e1, d.data(), &s));
      tot += s;
    });
    return std::sqrt(tot);
  }

  // calc_l2_error_gradient(mesh, fem_space, x, func): func(p, g) writes the
  // exact gradient at p into g[0..2]
  double calc_l2_error_gradient(const RegularMesh&, const Vector<T>& x, void (*func)(double*, T*)) {
    double tot = 0.0;
    for_chunks([&](int e0, int e1, const std::vector<double>& pts, std::vector<double>& val) {
      std::vector<double> g(3 * val.size());
      for (size_t q = 0; q < val.size(); ++q) {
        T gq[3];
        func(const_cast<double*>(&pts[3 * q]), gq);
        for (int d = 0; d < 3; ++d) g[3 * q + d] = (double)gq[d];
      }
      Vector<double> dg(g.size());
      dg.assign(g);
      double s = 0;
      check(tsem_h1_error_chunk(sp_, x.data(), e0, e1, dg.data(), &s));
      tot += s;
    });
    return std::sqrt(tot);
  }

  // calc_val_point(mesh, pos, u): u_h at the point pos[0..2] (NaN outside
  // the mesh)
  T calc_val_point(const RegularMesh&, const double* pos, const Vector<T>& u) {
    double out = 0.0;
    check(tsem_point_values(sp_, u.data(), 1, pos, &out, nullptr));
    return (T)out;
  }

  // write_coef(src, filename) / read_coef(dst, filename): 'M<TAB>n_dof', then
  // one value per line (14 significant digits); read_coef takes, per
  // geometry, the leading coefficients of a file of another order
  // (src/Miscellaneous.cpp:34-125)
  void write_coef(const std::vector<Vector<T>>& src, const std::string& filename) const {
    std::ofstream out(filename);
    if (!out) throw std::runtime_error("cannot write " + filename);
    out << M_ << '\t' << n_dof_total << '\n';
    char buf[64];
    for (const auto& v : src)
      for (T t : v.host()) {
        std::snprintf(buf, sizeof buf, "%.14g\n", (double)t);
        out << buf;
      }
  }
  void write_coef(const Vector<T>& src, const std::string& filename) const {
    write_coef(std::vector<Vector<T>>{src}, filename);
  }
  void read_coef(std::vector<Vector<T>>& dst, const std::string& filename) const {
    std::ifstream in(filename);
    if (!in) throw std::runtime_error("cannot open " + filename);
    int m_in = 0;
    long n_in = 0;
    in >> m_in >> n_in;
    auto binom = [](int n, int k) {
      if (k < 0 || n < k) return 0L;
      long r = 1;
      for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
      return r;
    };
    int64_t info[12];
    check(tsem_space_info(sp_, info));
    const int ng[4] = {(int)info[2], (int)info[3], (int)info[4], (int)info[5]};
    std::vector<int64_t> start((size_t)ng[0] + ng[1] + ng[2] + ng[3]);
    check(tsem_space_dof_start(sp_, start.data()));
    std::vector<int> dim(start.size());
    for (size_t g = 0, o = 0; o < 4; ++o)
      for (int i = 0; i < ng[o]; ++i) dim[g++] = (int)o;
    std::vector<size_t> order(start.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return start[a] < start[b]; });
    for (auto& v : dst) {
      std::vector<T> src(n_in), out(n_dof_total, T(0));
      for (auto& t : src) in >> t;
      if (!in) throw std::runtime_error("truncated coefficient file " + filename);
      long pos_in = 0;
      for (size_t g : order) {
        const long nd_in = binom(m_in - 1, dim[g]), keep = std::min(nd_in, binom(M_ - 1, dim[g]));
        for (long i = 0; i < keep; ++i) out[start[g] + i] = src[pos_in + i];
        pos_in += nd_in;
      }
      v.assign(out);
    }
  }
  void read_coef(Vector<T>& dst, const std::string& filename) const {
    std::vector<Vector<T>> one(1);
    read_coef(one, filename);
    dst = one[0];
  }

  // calc_sum_dof: polynomial degree of every DOF
  void calc_sum_dof(std::vector<unsigned>& sum_dof) const {
    std::vector<int32_t> s(n_dof_total);
    check(tsem_space_sum_dof(sp_, s.data()));
    sum_dof.assign(s.begin(), s.end());
  }

 private:
  tsem_space* sp_ = nullptr;
  int M_ = 0, nt_ = 0, nq3_ = 0, qorder_ = 0;
  std::vector<int32_t> face_vtx_, mark_f_;

  void init(int M, int qorder, RegularMesh& mesh, Context& ctx) {
    M_ = M;
    qorder_ = qorder;
    check(tsem_space_create(ctx.get(), M, (int)(mesh.xyz.size() / 3), mesh.xyz.data(), (int)(mesh.tet.size() / 4),
                            mesh.tet.data(), nullptr, qorder, &sp_));
    int64_t info[12];
    check(tsem_space_info(sp_, info));
    n_dof_total = (int)info[0];
    nt_ = (int)info[5];
    nq3_ = (int)info[11];
    stiff_matrix = SparseMatrix<T>(sp_, 1.0, 0.0);
    mass_matrix = SparseMatrix<T>(sp_, 0.0, 1.0);
  }
  void bc_points(int flag, std::vector<double>& pv, std::vector<double>& pe, std::vector<double>& pf) {
    int64_t n3[3];
    check(tsem_bc_count(sp_, flag, n3));
    pv.resize(3 * (size_t)n3[0]);
    pe.resize(3 * (size_t)n3[1] * qorder_);
    pf.resize(3 * (size_t)n3[2] * qorder_ * qorder_);
    check(tsem_bc_points(sp_, flag, pv.data(), pe.data(), pf.data()));
  }
  template <typename F>
  void for_chunks(F&& body) {
    const int chunk = std::max(1, std::min(nt_, (int)(4e7 / std::max(1, nq3_ * 3))));
    for (int e0 = 0; e0 < nt_; e0 += chunk) {
      int e1 = std::min(nt_, e0 + chunk);
      Vector<double> dpts(3 * (size_t)(e1 - e0) * nq3_);
      check(tsem_qp_coords(sp_, e0, e1, dpts.data()));
      std::vector<double> pts = dpts.host(), val((size_t)(e1 - e0) * nq3_);
      body(e0, e1, pts, val);
    }
  }
};

// ---------------------------------------------------------- TSEMSolver --
template <typename T>
class TSEMSolver {
 public:
  // init_PMG(A, mass, p, sum_dof): the device p-multigrid is built from the
  // space on first use; kept for call compatibility
  void init_PMG(const SparseMatrix<T>&, const SparseMatrix<T>&, int, std::vector<unsigned>&) {}

  // solve_PCG(A, x, b, tol, max_step, flag, ite_step, var): the reference's
  // stopping rule (||r||_1 < tol, absolute), flag 1 Jacobi or 4 p-multigrid
  // (var = smoothing steps); x is the initial guess and the solution.
  // Returns the iteration count.
  int solve_PCG(const SparseMatrix<T>& A, Vector<T>& x, const Vector<T>& b, double tol = 1.0e-8,
                unsigned max_step = 100000, unsigned flag = 1, unsigned = 1, double var = 1) const {
    require_stiffness(A);
    tsem_pcg_opts o{tol, 0, (int)max_step, (int)flag, 0, 1e-4, 100000, std::max(1, (int)var), 1};
    tsem_pcg_stats s{};
    int rc = tsem_pcg(A.space(), b.data(), x.data(), &o, &s);
    if (rc != TSEM_ENOCONV) check(rc);
    return s.iterations;
  }

  // solve_InversePower(x, A, B, flag, tol, max_step): lowest eigenpair of
  // A x = lambda B x; returns lambda
  double solve_InversePower(Vector<T>& x, const SparseMatrix<T>& A, const SparseMatrix<T>&, bool, double tol,
                            unsigned max_step) const {
    require_stiffness(A);
    tsem_pcg_opts o{tol * 1.0e-2, 1, 100000, 1, 0, 1e-4, 100000, 2, 10};
    double lam = 0;
    int its = 0;
    int rc = tsem_inverse_power(A.space(), x.data(), tol, (int)max_step, &o, &lam, &its);
    if (rc != TSEM_ENOCONV) check(rc);
    return lam;
  }

  // solve_LOBPCG(tsem, x, A, B, flag, tol, max_step): the x.size() lowest
  // eigenpairs (B-orthonormal x), stopping when no eigenvalue changes by
  // tol; returns the eigenvalues
  std::vector<double> solve_LOBPCG(TSEM<T>& tsem, std::vector<Vector<T>>& x, const SparseMatrix<T>& A,
                                   const SparseMatrix<T>& B, bool, double tol, unsigned max_step) const {
    const int m = (int)x.size();
    const size_t n = (size_t)tsem.n_dof_total;
    Vector<T> X(n * m);
    for (int j = 0; j < m; ++j) Vector<T>::copy_device(X.data() + j * n, x[j].data(), n);
    tsem_lobpcg_opts o{tol, 0, (int)max_step, 4, 2, A.stiff(), A.mass(), 0.0};
    tsem_lobpcg_stats s{};
    std::vector<double> lam(m);
    (void)B;
    int rc = tsem_lobpcg(tsem.space(), m, X.data(), lam.data(), &o, &s);
    if (rc != TSEM_ENOCONV) check(rc);
    for (int j = 0; j < m; ++j) Vector<T>::copy_device(x[j].data(), X.data() + j * n, n);
    return lam;
  }

 private:
  static void require_stiffness(const SparseMatrix<T>& A) {
    if (A.stiff() != 1.0 || A.mass() != 0.0)
      throw std::invalid_argument("solve_PCG / solve_InversePower operate on the stiffness operator");
  }
};

}  // namespace tsem

#endif
