// C entry points of the CPU oracle for ctypes (TEST INFRASTRUCTURE ONLY).
#include <cmath>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include <omp.h>

#include "oracle_core.h"
#include "oracle_mg.h"
#include "oracle_solver.h"

using namespace oracle;

static thread_local std::string g_err;

extern "C" {

// OpenMP threads the oracle's parallel loops use (the CPU baseline's core count)
int oracle_num_threads(void) { return omp_get_max_threads(); }

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

void oracle_laplace_rhs_divergence_fourier(void * h, double const * modes, int n_modes, double t,
                                           double * b)
{
  static_cast<LaplaceOperator *>(h)->rhs_divergence_fourier(modes, n_modes, t, b);
}

void oracle_divergence_apply(void * h, int degree_u, double const * u, double * b)
{
  divergence_apply(*static_cast<LaplaceOperator *>(h), degree_u, u, b);
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

// the cells' diagonal blocks (n_cells x n^3 x n^3, column j of cell c at (c n^3 + j) n^3)
void oracle_laplace_block_diagonal(void * h, double * blocks)
{
  static_cast<LaplaceOperator *>(h)->block_diagonal(blocks);
}

// PCG preconditioned by block Jacobi (Cholesky-factorized cell blocks)
int oracle_cg_solve_block_jacobi(void * h, double const * b, double * x, double rel_tol,
                                 double abs_tol, int max_iter, double * info)
{
  auto * op = static_cast<LaplaceOperator *>(h);
  try
  {
    BlockJacobi   bj(*op);
    int64_t const n = op->n_dofs();
    Apply         A = [&](double * dst, double const * src) { op->vmult(dst, src); };
    Apply         P = [&](double * dst, double const * src) { bj.vmult(dst, src); };
    SolveInfo     si = pcg(n, A, &P, b, x, rel_tol, abs_tol, max_iter);
    info[0]          = si.iterations;
    info[1]          = si.r0;
    info[2]          = si.r_final;
    info[3]          = si.converged ? 1.0 : 0.0;
    return 0;
  }
  catch(std::exception const & e)
  {
    g_err = e.what();
    return 1;
  }
}

// hybrid multigrid (p-sequence: degrees fine to coarse; type: MGType; n_h_levels meshes;
// smoother preconditioner 0 point Jacobi, 1 block Jacobi, 2 additive Schwarz)
void * oracle_mg_create(oracle_mesh const * m, int const * degrees, int n_degrees, int coarse_q1,
                        int smoother_degree, double smoothing_range, int eig_its,
                        int coarse_solver, double coarse_rel_tol, int coarse_max_iter,
                        int singular, double ip_factor, int type, int n_h_levels,
                        int smoother_preconditioner)
{
  try
  {
    MeshDesc d;
    std::memcpy(&d, m, sizeof(MeshDesc));
    MGDesc md;
    md.degrees.assign(degrees, degrees + n_degrees);
    md.coarse_q1       = coarse_q1 != 0;
    md.type            = type;
    md.n_h_levels      = n_h_levels;
    md.smoother_preconditioner = smoother_preconditioner;
    md.smoother_degree = smoother_degree;
    md.smoothing_range = smoothing_range;
    md.eig_its         = eig_its;
    md.coarse_solver   = coarse_solver;
    md.coarse_rel_tol  = coarse_rel_tol;
    md.coarse_max_iter = coarse_max_iter;
    md.singular        = singular != 0;
    return new Multigrid(d, md, ip_factor);
  }
  catch(std::exception const & e)
  {
    g_err = e.what();
    return nullptr;
  }
}

void oracle_mg_destroy(void * h) { delete static_cast<Multigrid *>(h); }

void oracle_mg_vmult(void * h, double * dst, double const * src)
{
  static_cast<Multigrid *>(h)->vmult(dst, src);
}

// level list of a multigrid type (coarse to fine): h, degree, dg per level; returns the count
// (-1 and the message on an invalid combination)
int oracle_mg_level_list(int type, int const * degrees, int n_degrees, int n_h_levels,
                         int coarse_q1, int * h, int * degree, int * dg, int max_levels)
{
  try
  {
    std::vector<LevelInfo> const L =
      mg_level_list(type, std::vector<int>(degrees, degrees + n_degrees), n_h_levels, coarse_q1 != 0);
    for(size_t l = 0; l < L.size() && (int)l < max_levels; ++l)
    {
      h[l]      = L[l].h;
      degree[l] = L[l].degree;
      dg[l]     = L[l].dg ? 1 : 0;
    }
    return (int)L.size();
  }
  catch(std::exception const & e)
  {
    g_err = e.what();
    return -1;
  }
}

// per level (0 = coarsest): n, degree, dg flag, lambda_max estimate, Chebyshev degree, h
void oracle_mg_info(void * h, int * n_levels, long long * n, int * degree, int * dg, double * lmax,
                    int * cheb_degree, int * coarse_its, int * h_level)
{
  auto * mg = static_cast<Multigrid *>(h);
  *n_levels = (int)mg->levels.size();
  for(size_t l = 0; l < mg->levels.size(); ++l)
  {
    auto const & L   = mg->levels[l];
    n[l]             = L.n;
    h_level[l]       = L.h;
    degree[l]        = L.degree;
    dg[l]            = L.dg ? 1 : 0;
    Chebyshev const & c = l == 0 ? mg->coarse_cheb : L.smoother;
    lmax[l]          = c.lambda_max;
    cheb_degree[l]   = c.degree;
  }
  *coarse_its = mg->last_coarse_its;
}

void oracle_mg_level_vmult(void * h, int l, double * dst, double const * src)
{
  static_cast<Multigrid *>(h)->levels[l].A(dst, src);
}

void oracle_mg_prolongate_add(void * h, int l, double * fine, double const * coarse)
{
  static_cast<Multigrid *>(h)->prolongate_add(l, fine, coarse);
}

void oracle_mg_restrict_add(void * h, int l, double * coarse, double const * fine)
{
  static_cast<Multigrid *>(h)->restrict_add(l, coarse, fine);
}

// Chebyshev smoother of level l (l >= 1): dst = S src (vmult, zero initial guess); theta/delta
void oracle_mg_smoother_vmult(void * h, int l, double * dst, double const * src, double * theta_delta)
{
  auto const & L = static_cast<Multigrid *>(h)->levels[l];
  std::vector<double> b(src, src + L.n), x(L.n, 0.0);
  L.smoother.vmult(x, b);
  std::copy(x.begin(), x.end(), dst);
  theta_delta[0] = L.smoother.theta;
  theta_delta[1] = L.smoother.delta;
}

// PCG on the SIPG operator with the multigrid preconditioner; rel_mode 0: vs ||r0||, 1: vs ||b||
int oracle_cg_solve_mg(void * h_op, void * h_mg, double const * b, double * x, double rel_tol,
                       double abs_tol, int max_iter, int rel_mode, double * info)
{
  auto * op = static_cast<LaplaceOperator *>(h_op);
  auto * mg = static_cast<Multigrid *>(h_mg);
  try
  {
    int64_t const n = op->n_dofs();
    Apply         A = [&](double * dst, double const * src) { op->vmult(dst, src); };
    Apply         P = [&](double * dst, double const * src) { mg->vmult(dst, src); };
    double        tol_rel = rel_tol;
    if(rel_mode == 1)
    {
      // relative to ||b||: convert to an absolute tolerance
      double bn = 0;
      for(int64_t i = 0; i < n; ++i)
        bn += b[i] * b[i];
      abs_tol = std::max(abs_tol, rel_tol * std::sqrt(bn));
      tol_rel = 0.0;
    }
    SolveInfo si = pcg(n, A, mg ? &P : nullptr, b, x, tol_rel, abs_tol, max_iter);
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
