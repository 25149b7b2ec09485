// Laplace eigenvalue problem on [0,1]^3 with zero Dirichlet data, the call
// sequence of examples/laplaceEig/laplaceEig.cpp:70-100 on the device
// library: lowest eigenpair by inverse power iteration from a random start,
// eigenvalue as the Rayleigh quotient against 3 pi^2.  With a third argument
// m > 1 the m lowest pairs are computed by block LOBPCG instead.
//
//   laplace_eig mesh_file polynomial_order [m]
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>

#include "tsem.hpp"

static const double PI = 3.14159265358979323846;

int main(int argc, char* argv[]) {
  if (argc < 3) {
    std::printf("Usage: %s filename_mesh polynomial_order [m]\n", argv[0]);
    return 1;
  }
  try {
    tsem::RegularMesh mesh;
    mesh.readMesh(argv[1]);
    const int M = std::atoi(argv[2]);
    const int m = argc > 3 ? std::atoi(argv[3]) : 1;

    tsem::TSEM<double> tsem(M, mesh);
    tsem.build_flag_bm(mesh);
    const int n_dof_total = tsem.n_dof_total;

    tsem::SparseMatrix<double> stiff(tsem.stiff_matrix.get_sparsity_pattern());
    stiff.copy_from(tsem.stiff_matrix);
    tsem.impose_zero_boundary_condition(stiff);
    tsem::SparseMatrix<double> mass(tsem.mass_matrix.get_sparsity_pattern());
    mass.copy_from(tsem.mass_matrix);
    tsem.impose_zero_boundary_condition(mass);

    std::mt19937_64 gen(7);
    std::uniform_real_distribution<double> uni(0.0, 1.0);
    tsem::TSEMSolver<double> solver;
    const double eig_ref = 3 * PI * PI;
    if (m <= 1) {
      std::vector<double> h(n_dof_total);
      for (auto& v : h) v = uni(gen);
      tsem::Vector<double> solution(n_dof_total);
      solution.assign(h);
      tsem.impose_zero_boundary_condition(solution);
      solver.solve_InversePower(solution, stiff, mass, false, 1.0e-12, 1000);

      // Rayleigh quotient with the unconstrained operators (laplaceEig.cpp:89-95)
      tsem::Vector<double> tmpv(n_dof_total);
      tsem.stiff_matrix.vmult(tmpv, solution);
      const double numerator = tmpv * solution;
      tsem.mass_matrix.vmult(tmpv, solution);
      const double denominator = tmpv * solution;
      const double eig = numerator / denominator;
      std::printf("n_element = %d, M = %d, n_dof = %d, eig = %.16e, err_eig = %.6e\n", mesh.n_geometry(3), M,
                  n_dof_total, eig, std::fabs(eig - eig_ref));
    } else {
      std::vector<tsem::Vector<double>> x(m, tsem::Vector<double>(n_dof_total));
      for (auto& xi : x) {
        std::vector<double> h(n_dof_total);
        for (auto& v : h) v = uni(gen) - 0.5;
        xi.assign(h);
      }
      std::vector<double> lam = solver.solve_LOBPCG(tsem, x, stiff, mass, false, 1.0e-10, 1000);
      std::printf("n_element = %d, M = %d, n_dof = %d, eig =", mesh.n_geometry(3), M, n_dof_total);
      for (double l : lam) std::printf(" %.16e", l);
      std::printf(", err_eig = %.6e\n", std::fabs(lam[0] - eig_ref));
    }
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 2;
  }
  return 0;
}
