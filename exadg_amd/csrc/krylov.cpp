#include "krylov.h"

#include <cmath>

namespace ExaDG
{
namespace Krylov
{
SolveInfo SolverCG::solve(Poisson::LaplaceOperator<double> const & A, PreconditionerBase * P,
                          double const * b, double * x, SolverData const & data, hipStream_t s)
{
  if(b == nullptr || x == nullptr)
    EXADG_THROW(EXADG_ERR_ARGUMENT, "null vector");
  if(b == x)
    EXADG_THROW(EXADG_ERR_ARGUMENT, "b and x must not alias");
  int64_t const n = A.m();
  ws_.reinit(n);
  double *  r = ws_.r.data(), *z = P ? ws_.z.data() : ws_.r.data(), *p = ws_.p.data(),
         *  v = ws_.v.data();
  SolveInfo info;

  // r = b - A x, z = P r
  A.residual(r, x, b, s);
  if(P)
    P->vmult(z, r, s);
  vec::norm2_and_dot(red_, 1, 2, r, z, n, s);
  vec::dot(red_, 4, b, b, n, s);
  double const * h = red_.fetch(s, 5);
  info.r0          = std::sqrt(h[1]);
  info.b_norm      = std::sqrt(h[4]);
  info.r_final     = info.r0;
  if(!std::isfinite(info.r0))
  {
    info.not_finite = true;
    return info;
  }
  double const tol = std::max(data.abs_tol, data.rel_tol * (data.rel_mode ? info.b_norm : info.r0));
  info.tolerance   = tol;
  info.r_true      = info.r0;
  if(info.r0 <= tol)
  {
    info.converged = true;
    return info;
  }
  vec::copy(p, z, n, s);
  for(int it = 1; it <= data.max_iter; ++it)
  {
    int const old_slot = 2 + (it - 1) % 2, new_slot = 2 + it % 2;
    A.vmult(v, p, s);
    vec::dot(red_, 0, p, v, n, s);
    vec::cg_update(x, r, p, v, red_.result() + old_slot, red_.result(), n, s);
    if(P)
      P->vmult(z, r, s);
    vec::norm2_and_dot(red_, 1, new_slot, r, z, n, s);
    h               = red_.fetch(s, 4);
    double const rn = std::sqrt(h[1]);
    info.iterations = it;
    info.r_final    = rn;
    if(!std::isfinite(rn))
    {
      info.not_finite = true;
      break;
    }
    if(rn <= tol)
    {
      info.converged = true;
      break;
    }
    vec::cg_direction(p, z, red_.result() + new_slot, red_.result() + old_slot, n, s);
  }
  if(!info.not_finite)
  {
    // final true residual in FP64: one fused residual kernel and a norm
    A.residual(v, x, b, s);
    vec::dot(red_, 5, v, v, n, s);
    info.r_true = std::sqrt(red_.fetch(s, 6)[5]);
    if(!std::isfinite(info.r_true))
      info.not_finite = true;
    else
      info.true_residual_ok = info.r_true <= 10.0 * tol;
  }
  if(info.iterations > 0 && info.r0 > 0 && info.r_final > 0)
  {
    info.rho = std::pow(info.r_final / info.r0, 1.0 / info.iterations);
    info.n10 = (info.rho < 1.0 && info.rho > 0) ? -10.0 * std::log(10.0) / std::log(info.rho) : 0.0;
  }
  return info;
}
}  // namespace Krylov
}  // namespace ExaDG
