// CPU oracle solvers (TEST INFRASTRUCTURE ONLY). See oracle_solver.h for the reference files.
#include "oracle_solver.h"

#include <cmath>
#include <stdexcept>

namespace oracle
{
static double dot(int64_t n, double const * a, double const * b)
{
  double s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for(int64_t i = 0; i < n; ++i)
    s += a[i] * b[i];
  return s;
}

SolveInfo pcg(int64_t n, Apply const & A, Apply const * P, double const * b, double * x,
              double rel_tol, double abs_tol, int max_iter)
{
  std::vector<double> r(n), z(n), p(n), v(n);
  A(v.data(), x);
#pragma omp parallel for schedule(static)
  for(int64_t i = 0; i < n; ++i)
    r[i] = b[i] - v[i];
  SolveInfo info;
  double    rn = std::sqrt(dot(n, r.data(), r.data()));
  info.r0      = rn;
  double const tol = std::max(abs_tol, rel_tol * rn);
  if(!std::isfinite(rn))
    throw std::runtime_error("non-finite initial residual");
  if(rn <= tol)
  {
    info.r_final   = rn;
    info.converged = true;
    return info;
  }
  if(P)
    (*P)(z.data(), r.data());
  else
    z = r;
  p          = z;
  double rho = dot(n, r.data(), z.data());
  for(int it = 1; it <= max_iter; ++it)
  {
    A(v.data(), p.data());
    double const alpha = rho / dot(n, p.data(), v.data());
#pragma omp parallel for schedule(static)
    for(int64_t i = 0; i < n; ++i)
    {
      x[i] += alpha * p[i];
      r[i] -= alpha * v[i];
    }
    rn              = std::sqrt(dot(n, r.data(), r.data()));
    info.iterations = it;
    if(!std::isfinite(rn))
      throw std::runtime_error("non-finite residual in CG");
    if(rn <= tol)
    {
      info.converged = true;
      break;
    }
    if(P)
      (*P)(z.data(), r.data());
    else
      z = r;
    double const rho_new = dot(n, r.data(), z.data());
    double const beta    = rho_new / rho;
    rho                  = rho_new;
#pragma omp parallel for schedule(static)
    for(int64_t i = 0; i < n; ++i)
      p[i] = z[i] + beta * p[i];
  }
  info.r_final = rn;
  return info;
}

void invert_diagonal(std::vector<double> & d)
{
  for(auto & v : d)
    v = (std::fabs(v) > 1e-10) ? 1.0 / v : 1.0;
}

JacobiCG::JacobiCG(LaplaceOperator const & op_, bool use_jacobi_) : op(op_), use_jacobi(use_jacobi_)
{
  if(use_jacobi)
  {
    inv_diag.resize(op.n_dofs());
    op.diagonal(inv_diag.data());
    invert_diagonal(inv_diag);
  }
}

SolveInfo JacobiCG::solve(double const * b, double * x, double rel_tol, double abs_tol,
                          int max_iter)
{
  int64_t const n = op.n_dofs();
  Apply A = [&](double * dst, double const * src) { op.vmult(dst, src); };
  Apply P = [&](double * dst, double const * src) {
#pragma omp parallel for schedule(static)
    for(int64_t i = 0; i < n; ++i)
      dst[i] = inv_diag[i] * src[i];
  };
  return pcg(n, A, use_jacobi ? &P : nullptr, b, x, rel_tol, abs_tol, max_iter);
}

}  // namespace oracle
