// Poisson problem on [0,1]^3 with mixed boundary conditions, the call
// sequence of examples/poisson_mixBC/poisson.cpp:74-101 on the device
// library: Dirichlet data u on mark 1, Neumann data g on mark 2, PCG with the
// reference's absolute l1 stopping rule, L2 error against u.
//
//   poisson_mixbc mesh_file polynomial_order [tol] [precond] [coef_file]
//
// precond: 1 Jacobi (the reference's default), 4 p-multigrid.
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "tsem.hpp"

// every member of the interface's class templates is compiled here
template class tsem::Vector<double>;
template class tsem::SparseMatrix<double>;
template class tsem::TSEM<double>;
template class tsem::TSEMSolver<double>;

static const double PI = 3.14159265358979323846;

// the exact solution and its data (poisson.cpp:25-27)
double u(double* p) { return std::cos(PI * p[0]) * std::sin(1.5 * PI * p[1]) * std::sin(0.25 * PI + 2 * PI * p[2]); }
double f(double* p) { return 7.25 * PI * PI * u(p); }
double g(double* p) {
  return -2 * PI * std::cos(PI * p[0]) * std::sin(1.5 * PI * p[1]) * std::cos(0.25 * PI + 2 * PI * p[2]);
}
// grad u (poisson.cpp:28-33)
void grad_u(double* p, double* dst) {
  dst[0] = -PI * std::sin(PI * p[0]) * std::sin(1.5 * PI * p[1]) * std::sin(0.25 * PI + 2 * PI * p[2]);
  dst[1] = 1.5 * PI * std::cos(PI * p[0]) * std::cos(1.5 * PI * p[1]) * std::sin(0.25 * PI + 2 * PI * p[2]);
  dst[2] = 2 * PI * std::cos(PI * p[0]) * std::sin(1.5 * PI * p[1]) * std::cos(0.25 * PI + 2 * PI * p[2]);
}

int main(int argc, char* argv[]) {
  if (argc < 3) {
    std::printf("Usage: %s filename_mesh polynomial_order [tol] [precond] [coef_file]\n", argv[0]);
    return 1;
  }
  try {
    tsem::RegularMesh mesh;
    mesh.readMesh(argv[1]);
    const int M = std::atoi(argv[2]);
    const double tol = argc > 3 ? std::atof(argv[3]) : 1.0e-14;
    const unsigned precond = argc > 4 ? (unsigned)std::atoi(argv[4]) : 1;

    tsem::TSEM<double> tsem(M, 30, mesh);
    tsem.build_flag_bm(mesh);
    const int n_dof_total = tsem.n_dof_total;

    tsem::SparseMatrix<double> sp_matrix(tsem.stiff_matrix.get_sparsity_pattern());
    sp_matrix.copy_from(tsem.stiff_matrix);

    tsem::Vector<double> rhs(n_dof_total);
    tsem.calc_rhs(rhs, &f);
    tsem.impose_boundary_condition(rhs, mesh, &g, 2);
    tsem.impose_boundary_condition(sp_matrix, rhs, mesh, &u, true, 1);

    tsem::TSEMSolver<double> solver;
    tsem::Vector<double> solution(n_dof_total);
    const int its = solver.solve_PCG(sp_matrix, solution, rhs, tol, 100000, precond, 1, 2);

    const double err_l2 = tsem.calc_l2_error(solution, &u);
    const double err_l2g = tsem.calc_l2_error_gradient(mesh, solution, &grad_u);
    double pos[3] = {0.3, 0.4, 0.5};
    const double u_pt = tsem.calc_val_point(mesh, pos, solution);
    std::printf("n_element = %d, M = %d, n_dof = %d, iterations = %d, err_l2 = %.16e, err_l2g = %.16e, "
                "u_h(0.3,0.4,0.5) = %.16e, u(0.3,0.4,0.5) = %.16e\n",
                mesh.n_geometry(3), M, n_dof_total, its, err_l2, err_l2g, u_pt, u(pos));
    if (argc > 5) tsem.write_coef(solution, argv[5]);  // optional: the solution's coefficients
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 2;
  }
  return 0;
}
