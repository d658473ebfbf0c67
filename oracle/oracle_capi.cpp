// C entry points of the CPU oracle for ctypes (TEST INFRASTRUCTURE ONLY).
#include <cmath>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "oracle_core.h"
#include "oracle_solver.h"

using namespace oracle;

static thread_local std::string g_err;

extern "C" {

struct oracle_mesh
{
  int    n_cells[3];
  double lower[3];
  double upper[3];
  double deformation;
  int    frequency;
  int    mapping_degree;
  int    periodic[3];
  int    boundary_type[6];
};

char const * oracle_last_error() { return g_err.c_str(); }

void * oracle_laplace_create(oracle_mesh const * m, int degree, double ip_factor)
{
  try
  {
    MeshDesc d;
    std::memcpy(&d, m, sizeof(MeshDesc));
    return new LaplaceOperator(d, degree, ip_factor);
  }
  catch(std::exception const & e)
  {
    g_err = e.what();
    return nullptr;
  }
}

void oracle_laplace_destroy(void * h) { delete static_cast<LaplaceOperator *>(h); }

long long oracle_laplace_n_dofs(void * h) { return static_cast<LaplaceOperator *>(h)->n_dofs(); }

void oracle_laplace_vmult(void * h, double * dst, double const * src)
{
  static_cast<LaplaceOperator *>(h)->vmult(dst, src);
}

void oracle_laplace_rhs_sine(void * h, double omega, double * b)
{
  static_cast<LaplaceOperator *>(h)->rhs_sine(omega, b);
}

double oracle_laplace_error_sine(void * h, double omega, double const * u)
{
  return static_cast<LaplaceOperator *>(h)->error_sine(omega, u);
}

void oracle_laplace_diagonal(void * h, double * d) { static_cast<LaplaceOperator *>(h)->diagonal(d); }

void oracle_laplace_tau(void * h, double * tau)
{
  auto * op = static_cast<LaplaceOperator *>(h);
  std::memcpy(tau, op->tau_cell.data(), sizeof(double) * op->tau_cell.size());
}

// cell geometry of cell c at the n^3 Gauss points: jinv (9 per point), jxw
void oracle_laplace_cell_geometry(void * h, long long c, double * jinv, double * jxw)
{
  auto *    op = static_cast<LaplaceOperator *>(h);
  int const n3 = op->n3;
  std::memcpy(jinv, &op->cell_jinv[(size_t)c * n3 * 9], sizeof(double) * n3 * 9);
  std::memcpy(jxw, &op->cell_jxw[(size_t)c * n3], sizeof(double) * n3);
}

void oracle_laplace_face_geometry(void * h, long long c, int dir, int side, double * jinv,
                                  double * jxw, double * normal, double * xq)
{
  static_cast<LaplaceOperator *>(h)->face_geometry(c, dir, side, jinv, jxw, normal, xq);
}

void oracle_tables(int degree, double * gl, double * gp, double * gw, double * S, double * Sd,
                   double * d0, double * d1)
{
  Tables1D t = make_tables(degree);
  int const n = t.n;
  std::memcpy(gl, t.gl.data(), n * sizeof(double));
  std::memcpy(gp, t.gp.data(), n * sizeof(double));
  std::memcpy(gw, t.gw.data(), n * sizeof(double));
  std::memcpy(S, t.S.data(), n * n * sizeof(double));
  std::memcpy(Sd, t.Sd.data(), n * n * sizeof(double));
  std::memcpy(d0, t.d0.data(), n * sizeof(double));
  std::memcpy(d1, t.d1.data(), n * sizeof(double));
}

void oracle_random_vector(long long n, unsigned long long seed, double * v)
{
#pragma omp parallel for schedule(static)
  for(long long i = 0; i < n; ++i)
    v[i] = splitmix_uniform(seed, (uint64_t)i);
}

// PCG (Jacobi or none) to ||r|| <= max(abs_tol, rel_tol*||r0||) (ReductionControl semantics,
// iterative_solvers_dealii_wrapper.h:136-221). info: [iterations, ||r0||, ||r_final||, converged]
int oracle_cg_solve(void * h, int use_jacobi, double const * b, double * x, double rel_tol,
                    double abs_tol, int max_iter, double * info)
{
  auto * op = static_cast<LaplaceOperator *>(h);
  try
  {
    JacobiCG solver(*op, use_jacobi != 0);
    SolveInfo si = solver.solve(b, x, rel_tol, abs_tol, max_iter);
    info[0]      = si.iterations;
    info[1]      = si.r0;
    info[2]      = si.r_final;
    info[3]      = si.converged ? 1.0 : 0.0;
    return 0;
  }
  catch(std::exception const & e)
  {
    g_err = e.what();
    return 1;
  }
}

}  // extern "C"
