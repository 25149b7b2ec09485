/*
 * C ABI of the MI355X TSEM operator and solver library (libtsem_amd.so).
 *
 * Every entry point returns a status code (TSEM_OK on success); on failure
 * tsem_last_error() describes the cause.  Vectors marked "device" are HIP
 * device pointers of length n_dof (double or float as stated); all work runs
 * on the context's stream and the calls return after it completed unless
 * stated otherwise.  Device buffers passed in must be complete (no pending
 * writes on other streams): the library works on its own stream.  One
 * context per host thread; calls on one context are serialized.
 *
 * Each function names the TSEMLib interface it replaces (file:line in the
 * reference tree).
 */
#ifndef TSEM_AMD_H
#define TSEM_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  TSEM_OK = 0,
  TSEM_ENOCONV = 1, /* iteration limit reached before the tolerance */
  TSEM_EINVAL = 2,  /* invalid argument */
  TSEM_ENOMEM = 3,  /* device or host allocation failed */
  TSEM_EDEVICE = 4, /* HIP runtime error */
  TSEM_ESTATE = 5   /* call out of order (e.g. solve before imposing BCs) */
};

typedef struct tsem_ctx tsem_ctx;
typedef struct tsem_space tsem_space;

const char* tsem_last_error(void);

/* Context: device + stream.  No counterpart (TSEMLib is host-only). */
int tsem_ctx_create(int device, tsem_ctx** out);
int tsem_ctx_destroy(tsem_ctx* ctx);
/* Solve class of a context, for several contexts sharing one device: 0
 * normal; 1 foreground; 2 background: a background tsem_pcg waits between
 * its chunks of check_every iterations while a foreground tsem_pcg runs
 * anywhere in the process, each time for at most as long as its previous
 * chunk took (host-side cooperative scheduling; streams keep the default
 * priority).  No counterpart (serving only). */
int tsem_ctx_set_class(tsem_ctx* ctx, int cls);
/* Explicit foreground span (any context): counts as foreground work until its
 * end, e.g. around a whole request (setup and solve).  A background solve
 * waits before each chunk while foreground work is active, for at most the
 * duration of its previous chunk.  tsem_background_wait returns at once for
 * a context that is not background, else when no foreground work is active
 * or after max_seconds (<= 0: no limit): a wait point for host work of
 * background requests. */
int tsem_foreground_begin(tsem_ctx* ctx);
int tsem_foreground_end(tsem_ctx* ctx);
int tsem_background_wait(tsem_ctx* ctx, double max_seconds);

/* Space of polynomial order M on a tetrahedral mesh; builds the index maps,
 * reference matrices, geometric factors and the assembled diagonal.
 * Replaces TSEM<double>::TSEM(M, mesh, fem_space) + init
 * (include/TetrahedralSEM.h:193-215, src/Initialize.cpp:11-63).
 * tet: nt x 4 vertex ids (element vertex order P0..P3); femdof_weight:
 * optional (nv + n_edges) ordering weights of vertices then edges (the P2 fem
 * dof numbers AFEPack supplies), NULL = the canonical element-scan order;
 * qorder: Gauss-Legendre points per direction of the collapsed quadrature for
 * the load vector and boundary data (TSEM(M, qinfo_order, ...),
 * src/QuadInfo.cpp:63-119). */
int tsem_space_create(tsem_ctx* ctx, int M, int nv, const double* xyz, int nt, const int32_t* tet,
                      const int32_t* femdof_weight, int qorder, tsem_space** out);
int tsem_space_destroy(tsem_space* sp);
/* ctx may be NULL: a host-only space (setup, maps and reference data only;
 * device entry points then return TSEM_ESTATE). */

/* Replace the vertex coordinates of a space (same connectivity, nv x 3
 * host): volumes and geometric factors are recomputed, the assembled
 * diagonal is rebuilt on the device, and boundary eliminations and the
 * potential are cleared (marks stay).  Lets one space serve many jittered
 * meshes of one topology.  No counterpart (the reference rebuilds TSEM). */
int tsem_space_set_geometry(tsem_space* sp, const double* xyz);

/* Assembled diagonal of the stiffness operator, host copy (n_dof). */
int tsem_space_diagonal_host(const tsem_space* sp, double* diag);

/* info[0..11] = n_dof, n_index(K), nv, ne, nf, nt, n_dof_geometry[0..3],
 * n_map_terms, n_qp3 */
int tsem_space_info(const tsem_space* sp, int64_t* info);

/* Derived topology: edge_vtx[ne][2], face_vtx[nf][3], face_edge[nf][3] (edge
 * opposite face vertex j), tet_face[nt][4] (face opposite vertex k). */
int tsem_space_topology(const tsem_space* sp, int32_t* edge_vtx, int32_t* face_vtx, int32_t* face_edge,
                        int32_t* tet_face);

/* Gather maps (host arrays): local coefficient (e, i) = sum_{t in
 * [off[eK+i], off[eK+i+1])} val[t] * x[ind[t]].  Replaces the public
 * transform_{n,ind,val}_global2local (include/TetrahedralSEM.h:97-99). */
int tsem_space_maps(const tsem_space* sp, int32_t* off, int32_t* ind, double* val);

/* First DOF of every geometry (vertices, edges, faces, tets in that order):
 * location_actualdof[location_geometry[.]] (src/BuildSpace.cpp:940-945). */
int tsem_space_dof_start(const tsem_space* sp, int64_t* start);

/* Entity-slab state: out[0] = 1 when the FP32 full-order operator of this
 * space writes entity slabs (tables on the device, switch on, matrix-core
 * element kernel with room for the tables), out[1] = slab entries (0: the
 * mesh has no slab layout), out[2] = largest slab span of a 256-DOF chunk,
 * out[3] = terms per transposed face row.  No counterpart. */
int tsem_space_slab(const tsem_space* sp, int64_t* out);

/* Host: y (n_dof) = the scatter of element-major results yl (n_elem x K)
 * through the entity-slab tables of the space, with the device kernels'
 * slab entry function (a check of the slab layout against tsem_space_maps).
 * TSEM_ESTATE when the mesh has no slab layout.  No counterpart. */
int tsem_space_slab_apply_host(const tsem_space* sp, const double* yl, double* y);

/* Polynomial degree of each DOF: TSEM::calc_sum_dof (src/Miscellaneous.cpp:557). */
int tsem_space_sum_dof(const tsem_space* sp, int32_t* sum_dof);

/* Reference matrices (host, K x K row-major): mass, 6 pair stiffness matrices
 * (stiff_matrix_element_actual, src/Discretize.cpp:844-1294), per-element
 * pair weights w[nt][6] (2/(3V) x dihedral factor, src/Discretize.cpp:1355-1370). */
int tsem_space_ref_matrices(const tsem_space* sp, double* mass, double* stiff6, double* pair_w);

/* Factorizations the element kernels use (host): dfac[3][Kd][K] with
 * dfac_a^T dfac_b = grad2[3a+b] = int d_a phi_i d_b phi_j, Kd = C(M+2,3);
 * mfac[K][K] with mfac^T mfac = mass.  No counterpart (the reference
 * assembles instead). */
int tsem_space_factors(const tsem_space* sp, double* dfac, double* mfac, double* grad2);

/* Boundary marks per vertex, edge, face: TSEM::build_flag_bm(mesh)
 * (src/Discretize.cpp:107-118). */
int tsem_set_marks(tsem_space* sp, const int32_t* mark_v, const int32_t* mark_e, const int32_t* mark_f);

/* Boundary data are given on the geometries that carry a mark, in geometry
 * index order.  tsem_bc_count: numbers of vertices, edges, faces with
 * mark == flag.  tsem_bc_points: where the data are sampled (host):
 * pv[nv_f][3]; pe[ne_f][q][3] along the first-visit edge direction;
 * pf[nf_f][q*q][3] in the first-visit face order (src/Discretize.cpp:1797-1830). */
int tsem_bc_count(const tsem_space* sp, int flag, int64_t* counts3);
int tsem_bc_points(const tsem_space* sp, int flag, double* pv, double* pe, double* pf);

/* Load vector, device: b = sum_e G_e^T V_e B^T W f_e  (TSEM::calc_rhs with
 * values at quadrature points, src/Discretize.cpp:1396-1418).  Call
 * tsem_rhs_begin, then tsem_rhs_add for element ranges with f at the 3-D
 * quadrature points (device, (e1-e0) x n_qp3), then tsem_rhs_end(b). */
int tsem_qp_coords(tsem_space* sp, int e0, int e1, double* xyz_dev);
int tsem_rhs_begin(tsem_space* sp);
int tsem_rhs_add(tsem_space* sp, int e0, int e1, const double* fq_dev);
int tsem_rhs_end(tsem_space* sp, double* b_dev);

/* The manufactured Poisson family of the examples and benchmarks:
 * u = sin(pi x) sin(pi y) sin(pi z) (1 + amp g), g = 1/nterms sum_m
 * cos(2 pi k_m . x + phi_m), k_m integer in 0..3.  tsem_manufactured
 * evaluates, at n device points (pts: n x 3), what = 0: u; 1: f = -Lap u;
 * 2: du/dn with the unit normals (device, n x 3), into out (device, n),
 * enqueued on `stream` (a hipStream_t; NULL: the null stream) without
 * waiting.  No counterpart (the reference's example evaluates its own
 * functions on the host, examples/poisson_mixBC/poisson.cpp:25-27). */
typedef struct {
  int nterms;        /* 1..16 */
  double k[16][3];   /* wave vectors */
  double phi[16];    /* phases */
  double amp;
} tsem_manufactured_def;
int tsem_manufactured(const tsem_manufactured_def* def, int what, int64_t n, const double* pts, const double* normals,
                      double* out, void* stream);

/* Neumann load on faces with mark == flag; g at their quadrature points
 * (host, nf_f x q*q, tsem_bc_points order).  b device, updated.
 * TSEM::impose_boundary_condition(rhs, mesh, g, flag) (src/Discretize.cpp:2010-2063). */
int tsem_neumann(tsem_space* sp, const double* g_face_qp, int flag, double* b_dev);

/* Dirichlet imposition on geometries with mark == flag: modal interpolation of
 * the data (uv[nv_f], ue[ne_f][q], uf[nf_f][q*q] host, at tsem_bc_points), b
 * updated and the operator switched to the eliminated form (rows/columns of
 * the constrained DOFs replaced by the diagonal).  Optionally returns the
 * imposed modal values (bval host, n_dof, NaN on free DOFs).
 * TSEM::impose_boundary_condition(A, rhs, mesh, u, true, flag)
 * (src/Discretize.cpp:1778-1893). */
int tsem_dirichlet(tsem_space* sp, const double* uv, const double* ue, const double* uf, int flag, double* b_dev,
                   double* bval);
/* Zero Dirichlet elimination (operator only) for mark == flag:
 * impose_zero_boundary_condition(A, flag) (src/Discretize.cpp:1465-1502). */
int tsem_zero_bc(tsem_space* sp, int flag);
/* Zero the entries of a device vector on the DOFs of geometries with
 * mark == flag: impose_zero_boundary_condition(Vector&, flag)
 * (include/TetrahedralSEM.h:251, src/Discretize.cpp:1539-1551). */
int tsem_zero_rows(tsem_space* sp, int flag, double* x_dev);

/* Matrix-free operator, device: y = (stiff * S + mass * M) x with the
 * Dirichlet elimination when constrained != 0.  prec 64: double vectors,
 * 32: float vectors.  Replaces SparseMatrix<double>::vmult on the assembled
 * TSEM::stiff_matrix / mass_matrix (src/TSEMSolver.cpp:1842). */
int tsem_apply(tsem_space* sp, double stiff, double mass, int constrained, int prec, const void* x, void* y);

/* tsem_apply without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  The same kernels and
 * results as tsem_apply (bit for bit), with the element-local scratch
 * (gathered element tiles, element results) in one device-resident array of
 * 64 MiB per device, held until the process ends, instead of the space's
 * buffers.  Spaces whose scratch exceeds it (about 4 M element-mode entries
 * in FP64, 8 M in FP32) get TSEM_EINVAL.  Every call on one device shares
 * that array: calls that overlap on one device (several threads or streams)
 * race; the array keeps the last call's contents, and every entry a call
 * reads it first writes. */
int tsem_apply_noalloc(tsem_space* sp, double stiff, double mass, int constrained, int prec, const void* x, void* y);

/* Element-local coefficients, device: xl[nt][K] (K1 gather). */
int tsem_gather(tsem_space* sp, const double* x_dev, double* xl_dev);
/* Transposed gather, device: y = sum_e G_e^T yl_e (K1 scatter, fixed order). */
int tsem_scatter(tsem_space* sp, const double* yl_dev, double* y_dev);

/* Kernel timing with HIP events recorded on the launching stream around the
 * full-order operator kernels (process wide).  tsem_timing(1) resets and
 * starts, tsem_timing(0) stops; tsem_timing_read fills out[12] = {total ms,
 * launches} for: gather FP64, element GEMM FP64, scatter FP64, gather FP32,
 * element GEMM FP32, scatter FP32 (the scalar element path counts as element
 * GEMM; synchronizes the device).  No counterpart. */
int tsem_timing(int enable);
int tsem_timing_read(double* out);

/* Element-kernel selection for the operator (process wide): 0 = fastest
 * matrix-core path (default: fp16-split MFMA for FP32 vectors, FP64 MFMA
 * for FP64), 1 = scalar kernel, 2 = FP32/FP64 MFMA (no fp16 splitting).
 * No counterpart. */
int tsem_set_apply_path(int path);

/* PCG iterations replayed as captured HIP graphs, one chunk of check_every
 * iterations per graph (process wide; default on, TSEM_PCG_GRAPH=0 turns it
 * off; kernel timing turns it off).  No counterpart. */
int tsem_set_graphs(int enable);

/* LOBPCG's p-multigrid preconditioner as one V-cycle over the block of
 * residual columns per iteration, every kernel one launch over the columns
 * (process wide; default on, TSEM_BLOCK_VCYCLE=0 turns it off, and it needs
 * the fp16-split FP32 operator on every level).  Off: one V-cycle per
 * column; the two give the same bits.  No counterpart. */
int tsem_set_block_vcycle(int enable);

/* Element results written as entity slabs (process wide; default on,
 * TSEM_SLAB=0 turns it off): the FP32 matrix-core element kernel stores each
 * element's contribution to each of its global DOFs, already through the
 * face re-expression, in per-geometry rows, and the scatter adds a DOF's
 * rows in element order with no index codes or weights.  Used for the
 * full-order operator without potential term; the sums run in another order
 * than tsem_scatter's (same values to rounding).  No counterpart. */
int tsem_set_slab(int enable);

/* Assembled diagonal (device copy out, n_dof). */
int tsem_diagonal(tsem_space* sp, double* diag_dev);

typedef struct {
  double tol;          /* tolerance */
  int criterion;       /* 0: reference, ||r||_1 < tol (absolute); 1: ||b-Ax||_2/||b||_2 <= tol */
  int max_iter;        /* iteration limit (mixed: limit on reliable updates / refinement steps) */
  int precond;         /* 1: Jacobi (reference default); 4: p-multigrid */
  int mixed;           /* 0: FP64 PCG (criterion 0 or 1);
                          1: FP32 PCG with reliable updates: the FP32 correction is added to
                             the FP64 x and the FP64 true residual replaces the FP32 one each
                             time the FP32 residual has fallen by inner_tol, search direction
                             kept (criterion 1 only);
                          2: FP64 iterative refinement around restarted FP32 PCG solves
                             (criterion 1 only) */
  double inner_tol;    /* mixed 1: residual reduction between reliable updates (clamped to
                          [1e-6, 0.5]); mixed 2: relative reduction per inner solve (the last
                          one stops once the outer tolerance is in reach, with a margin of 3) */
  int inner_max;       /* mixed: FP32 iteration limit (mixed 1: in total; 2: per inner solve) */
  int smooth_steps;    /* p-MG smoother degree */
  int check_every;     /* host convergence check interval (iterations) */
} tsem_pcg_opts;

typedef struct {
  int iterations;      /* total (inner for mixed) PCG iterations */
  int outer;           /* reliable updates (mixed 1) or refinement steps (mixed 2) */
  double relres;       /* final true ||b-Ax||_2/||b||_2 (FP64) */
  double l1;           /* final ||r||_1 of the recursive residual */
  double seconds;      /* wall time of the solve (b and x resident) */
} tsem_pcg_stats;

/* PCG on the eliminated operator, device b/x (x: initial guess, solution).
 * TSEMSolver<double>::solve_PCG(A, x, b, tol, max_step, flag, ite_step, var)
 * (include/TSEMSolver.h:89-90, src/TSEMSolver.cpp:1775-1930). */
int tsem_pcg(tsem_space* sp, const double* b_dev, double* x_dev, const tsem_pcg_opts* opts, tsem_pcg_stats* stats);

typedef struct {
  double tol;          /* criterion 1: max_i ||A x_i - lambda_i B x_i||_2 / (|lambda_i| ||B x_i||_2) <= tol;
                          criterion 0 (reference): max_i |lambda_i - lambda_i(previous)| < tol */
  int criterion;
  int max_iter;        /* iteration limit */
  int precond;         /* 0: none; 1: Jacobi; 4: one FP32 p-multigrid V-cycle per column */
  int smooth_steps;    /* p-MG smoother degree */
  double stiff;        /* A = stiff S + mass M + potential M_V (stiff > 0); B = M */
  double mass;
  double potential;    /* coefficient of M_V (tsem_set_potential); 0 = none */
} tsem_lobpcg_opts;

typedef struct {
  int iterations;
  int applies;         /* FP64 operator applies (A and B) */
  double max_resid;    /* final max relative residual */
  double max_dlambda;  /* final max eigenvalue change */
  double seconds;
} tsem_lobpcg_stats;

/* Block LOBPCG for the m (1..32) lowest eigenpairs of A x = lambda B x on
 * the eliminated operators (set the Dirichlet marks with tsem_zero_bc
 * first).  X: device, n_dof x m column-major (column j at X + j n_dof);
 * in: initial block, out: B-orthonormal eigenvectors, zero on eliminated
 * DOFs.  lambda: host, m values ascending.  Gram matrices and the
 * Rayleigh-Ritz problem are FP64 on the full 3m x 3m trial space.
 * TSEMSolver::solve_LOBPCG(tsem, x, A, B, flag, tol, max_step)
 * (include/TSEMSolver.h:95-98, src/TSEMSolver.cpp:2224-2407; the reference
 * solves a fixed 3 x 3 problem for one pair, corrected here). */
int tsem_lobpcg(tsem_space* sp, int m, double* X_dev, double* lambda, const tsem_lobpcg_opts* opts,
                tsem_lobpcg_stats* stats);

/* Inverse power iteration for the lowest eigenpair of S x = lambda M x:
 * x <- S^-1 M x by tsem_pcg with `inner`, M-normalize, stop when the change
 * of x^T S x is below tol.  x: device, initial guess in, eigenvector out.
 * TSEMSolver::solve_InversePower(x, A, B, flag, tol, max_step)
 * (include/TSEMSolver.h:93-94, src/TSEMSolver.cpp:2041-2069). */
int tsem_inverse_power(tsem_space* sp, double* x_dev, double tol, int max_iter, const tsem_pcg_opts* inner,
                       double* lambda, int* iterations);

/* ---- work at the 3-D quadrature points (n_qp3 points per element, weights
 * summing to 1 over the element, tsem_qp_coords gives the points) ---- */

/* u_h at the points of elements [e0, e1): out[(e - e0) n_qp3 + q] (device).
 * TSEM::calc_val_qp_onElement (src/PointInfo.cpp:60-91). */
int tsem_qp_values(tsem_space* sp, const double* x_dev, int e0, int e1, double* out_dev);

/* rho[e][q] = sum_k occ[k] psi_k(x_q)^2 over all elements (device, nt x
 * n_qp3); psi: device, n_dof x nk column-major; occ: host, nk values.
 * TSEM::calc_density (src/PointInfo.cpp:2012-2022). */
int tsem_density(tsem_space* sp, int nk, const double* psi_dev, const double* occ, double* rho_dev);

/* out[v] = sum_e vol_e sum_q w_q f g m_v(x_q) over all elements, for the
 * moments m = {1, x, y, z, xx, xy, xz, yy, yz, zz} (nmom = 1 or 10); f, g:
 * device, nt x n_qp3 (g may be NULL).  Used for the energies and the
 * multipole data of the Kohn-Sham example (examples/KohnSham-Helium/
 * SCFIteration.cpp:70-101, 303-326).  Fixed-order sums. */
int tsem_integrate(tsem_space* sp, const double* f_dev, const double* g_dev, int nmom, double* out);

/* sum over elements [e0, e1) of vol_e sum_q w_q (u_h(x_q) - u_q)^2, u: device
 * values at the points of those elements; the L2 error is the square root
 * of the sum over all chunks.  TSEM::calc_l2_error (src/CalcError.cpp:20-126). */
int tsem_l2_error_chunk(tsem_space* sp, const double* x_dev, int e0, int e1, const double* u_dev, double* sum);
/* Gradient error part of elements [e0, e1): *sum = sum_e V_e sum_q w_q
 * |grad u_h - g|^2 with g the exact gradient at the points (device,
 * (e1-e0) x n_qp3 x 3).  calc_l2_error_gradient(mesh, fem_space, sol, func)
 * (include/TetrahedralSEM.h:276, src/CalcError.cpp:199-267); the error is
 * the square root of the sum over all chunks. */
int tsem_h1_error_chunk(tsem_space* sp, const double* x_dev, int e0, int e1, const double* g_dev, double* sum);

/* Values of a solution at n arbitrary points: out[p] = u_h(pts[p]) (host
 * arrays: pts n x 3, out n); elem[p] (host, may be NULL) is the element
 * that holds the point, -1 (with out NaN) for a point outside the mesh.
 * Points are located on the host (element boxes bucketed on a grid; of the
 * elements holding a point on a shared face, edge or vertex, the lowest
 * numbered one, where the field is continuous anyway), the basis is
 * evaluated there and the sums run on the device.
 * TSEM::calc_val_point(mesh, pos, u) (src/PointInfo.cpp:1143-1313). */
int tsem_point_values(tsem_space* sp, const double* x_dev, int64_t n, const double* pts, double* out, int32_t* elem);

/* LDA (unpolarized) at n points: Slater exchange + Perdew-Zunger 1981
 * correlation; vxc = d(rho eps_xc)/d rho, exc = rho eps_xc (device; either
 * output may be NULL).  The reference example uses VWN correlation
 * (examples/KohnSham-Helium/SCFIteration.cpp:192-230). */
int tsem_lda_pz(tsem_space* sp, int64_t n, const double* rho_dev, double* vxc_dev, double* exc_dev);

/* Potential V at the points (device, nt x n_qp3, copied; NULL clears) that
 * weights the operator M_V, (M_V)_ij = int V phi_i phi_j, applied
 * matrix-free (TSEM::build_mass_V_matrix, src/Discretize.cpp:667-843;
 * needs qorder <= 20). */
int tsem_set_potential(tsem_space* sp, const double* v_dev);

/* y = (stiff S + mass M + potential M_V) x, FP64 device vectors. */
int tsem_apply_potential(tsem_space* sp, double stiff, double mass, double potential, int constrained,
                         const double* x_dev, double* y_dev);

/* ---- served stream of Poisson and eigen requests ----------------------
 * The reference has no served form; these four entry points are the
 * boundary of a service that solves many independent problems on one
 * device.  A Poisson request is the mixed-BC problem of
 * examples/poisson_mixBC/poisson.cpp:74-89 (load vector, Neumann load on
 * mark 2, Dirichlet elimination on mark 1), solved by PCG to a relative
 * residual; an eigen request the lowest pairs of the Dirichlet Laplacian by
 * LOBPCG; a system request a PCG solve of the caller's own assembled
 * system (its space, b and x), e.g. the Hartree solve of a Kohn-Sham SCF
 * step (examples/KohnSham-Helium/SCFIteration.cpp:154-177).  Workers (one
 * context each) take the smallest waiting request of their lane by estimated
 * work (elements x modes) -- the first small_slots workers only small
 * requests, the others the smallest large one, else the smallest small one
 * -- as soon as its device
 * footprint fits the budget; spaces are cached per topology key and only
 * their coordinates replaced; with classes, small requests run foreground
 * and large ones background (tsem_ctx_set_class). */
typedef struct tsem_server tsem_server;

/* f at n device points (xyz_dev: n x 3) into out_dev, enqueued on `stream`
 * (a hipStream_t) or finished before returning; nonzero = failure.  Called
 * from a worker thread. */
typedef int (*tsem_load_fn)(void* user, const double* xyz_dev, int64_t n, double* out_dev, void* stream);
/* u at n host points (n x 3) into out (host) */
typedef int (*tsem_value_fn)(void* user, const double* xyz, int64_t n, double* out);
/* du/dn at n host points with the outward unit normals (n x 3) into out */
typedef int (*tsem_flux_fn)(void* user, const double* xyz, const double* normal, int64_t n, double* out);

typedef struct {
  int slots;          /* worker threads, one context (stream) each; 0: 4 */
  int small_slots;    /* workers that take only small requests; -1: slots / 2 */
  double small_cost;  /* a request is small when n_elem * C(M+3,3) <= small_cost; 0: 4e6 */
  int64_t hbm_budget; /* device bytes the spaces in use and cached may hold; 0: 70% of free */
  int classes;        /* small requests foreground, large ones background */
} tsem_serve_opts;

typedef struct {
  int64_t topology;       /* caller's key: requests with one nonzero key share connectivity, order
                             and quadrature, and reuse a cached space (each request's own labels and
                             Dirichlet / Neumann parts are applied to it); 0: none */
  int M, qorder;          /* order; quadrature order (0: M + 3) */
  int nv, nt;
  const double* xyz;      /* host, nv x 3 (copied by submit) */
  const int32_t* tet;     /* host, nt x 4 (copied by submit) */
  const int32_t* labels;  /* host, nv: bits of the boundary parts each vertex lies on (copied); NULL: none */
  int32_t dirichlet;      /* label bits of the Dirichlet parts (mark 1); a geometry lies on a part when
                             all its vertices do */
  int32_t neumann;        /* label bits of the Neumann parts (mark 2); Dirichlet wins */
  tsem_load_fn f;         /* required */
  tsem_value_fn u;        /* required when a Dirichlet part exists */
  tsem_flux_fn g;         /* required when a Neumann part exists */
  void* user;             /* passed to the callbacks */
  double tol;             /* relative residual ||b - Ax||_2 / ||b||_2 (0: 1e-10) */
  int mixed;              /* tsem_pcg_opts.mixed (inner_tol 1e-2) */
  int precond;            /* 1 Jacobi (0 means 1), 4 p-multigrid */
  int kind;               /* TSEM_REQ_POISSON, TSEM_REQ_EIGEN or TSEM_REQ_SYSTEM */
  /* Poisson: the manufactured problem (copied by submit) in place of the
   * callbacks f, u and g, evaluated on the device; NULL: the callbacks */
  const tsem_manufactured_def* manufactured;
  /* eigen: the nev (1..32) lowest pairs of S x = lambda M x with zero
   * Dirichlet data on the Dirichlet parts, LOBPCG to residual tol (0: 1e-8)
   * from an N(0, 1) block of the given seed, p-multigrid preconditioned;
   * the eigenvalues go to eigenvalues (host, nev, written by the worker),
   * the eigenvectors (n_dof x nev, column-major) to poll's x */
  int nev;
  uint64_t seed;
  double* eigenvalues;
  /* system: PCG on the caller's space (its current elimination) for b
   * (device) from the initial guess x (device, solution out), on that
   * space's context; the mesh fields are unused and the space must not be
   * used by the caller until the request has finished */
  tsem_space* space;
  const double* b_dev;
  double* x_dev;
} tsem_serve_request;

#define TSEM_REQ_POISSON 0
#define TSEM_REQ_EIGEN 1
#define TSEM_REQ_SYSTEM 2

typedef struct {
  int status;             /* TSEM_OK, TSEM_ENOCONV, or the request's error */
  int built;              /* 1 when the request built its space (no cached one) */
  int64_t n_dof;
  int iterations, outer;
  double relres;
  double queued_s;         /* submit to start */
  double space_s;          /* space built, or its coordinates replaced */
  double setup_s;          /* load vector and boundary data */
  double solve_s;          /* PCG and the copy of the solution */
  double total_s;          /* submit to completion */
} tsem_serve_result;

int tsem_serve_open(int device, const tsem_serve_opts* opts, tsem_server** out);
/* Queues a request; *id names it in tsem_serve_poll. */
int tsem_serve_submit(tsem_server* s, const tsem_serve_request* req, int64_t* id);
/* *done = 0 while request id runs (wait != 0 blocks until it has finished).
 * Once finished: *done = 1, res filled (n_dof included) and the request's
 * status returned (its message in tsem_last_error); with x (host: n_dof
 * doubles, n_dof x nev for an eigen request, unused for a system request)
 * the solution is copied there and the request released, with x NULL it is
 * kept for a later poll. */
int tsem_serve_poll(tsem_server* s, int64_t id, int wait, double* x, tsem_serve_result* res, int* done);
/* Solves every queued request, then frees the server (results not polled are dropped). */
int tsem_serve_close(tsem_server* s);

#ifdef __cplusplus
}
#endif
#endif
