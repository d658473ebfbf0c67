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
  ws_.reinit(A.m());
  return cg_solve_in(A, P, b, x, data, ws_.r.data(), ws_.z.data(), ws_.p.data(), ws_.v.data(), red_,
                     s);
}

SolveInfo cg_solve_in(Poisson::LaplaceOperator<double> const & A, PreconditionerBase * P,
                      double const * b, double * x, SolverData const & data, double * r, double * z,
                      double * p, double * v, Reducer & red, hipStream_t s)
{
  if(b == nullptr || x == nullptr)
    EXADG_THROW(EXADG_ERR_ARGUMENT, "null vector");
  if(b == x)
    EXADG_THROW(EXADG_ERR_ARGUMENT, "b and x must not alias");
  int64_t const n = A.m();
  if(!P)
    z = r;
  SolveInfo info;
  // A FP32 preconditioner is applied in its own precision: the update kernel also writes the
  // residual rounded to FP32 into the preconditioner's input, and the dot product and the
  // direction update read its FP32 result, widened exactly. The iterates are the same as with
  // the FP64 vmult of the preconditioner, without its two conversions.
  float * const r_narrow = P ? P->narrow_input() : nullptr;
  float const * z_narrow = nullptr;

  // r = b - A x, z = P r
  A.residual(r, x, b, s);
  if(r_narrow)
  {
    vec::convert<float, double>(r_narrow, r, n, s);
    z_narrow = P->vmult_narrow(s);
    vec::norm2_and_dot(red, 1, 2, r, z_narrow, n, s);
  }
  else
  {
    if(P)
      P->vmult(z, r, s);
    vec::norm2_and_dot(red, 1, 2, r, z, n, s);
  }
  vec::dot(red, 4, b, b, n, s);
  double const * h = red.fetch(s, 5);
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
  if(z_narrow)
    vec::convert<double, float>(p, z_narrow, n, s);
  else
    vec::copy(p, z, n, s);
  for(int it = 1; it <= data.max_iter; ++it)
  {
    int const old_slot = 2 + (it - 1) % 2, new_slot = 2 + it % 2;
    A.vmult(v, p, s);
    vec::dot(red, 0, p, v, n, s);
    if(r_narrow)
    {
      vec::cg_update_narrow(x, r, r_narrow, p, v, red.result() + old_slot, red.result(), n, s);
      z_narrow = P->vmult_narrow(s);
      vec::norm2_and_dot(red, 1, new_slot, r, z_narrow, n, s);
    }
    else
    {
      vec::cg_update(x, r, p, v, red.result() + old_slot, red.result(), n, s);
      if(P)
        P->vmult(z, r, s);
      vec::norm2_and_dot(red, 1, new_slot, r, z, n, s);
    }
    h               = red.fetch(s, 4);
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
    if(z_narrow)
      vec::cg_direction(p, z_narrow, red.result() + new_slot, red.result() + old_slot, n, s);
    else
      vec::cg_direction(p, z, red.result() + new_slot, red.result() + old_slot, n, s);
  }
  if(!info.not_finite)
  {
    // final true residual in FP64: one fused residual kernel and a norm
    A.residual(v, x, b, s);
    vec::dot(red, 5, v, v, n, s);
    info.r_true = std::sqrt(red.fetch(s, 6)[5]);
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
