// extern "C" boundary: every entry point catches, records the message and returns a status.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>

#include "../../include/exadg_amd.h"
#include "common.h"
#include "divergence_operator.h"
#include "guess.h"
#include "krylov.h"
#include "laplace_operator.h"
#include "multigrid.h"
#include "noalloc.h"

namespace
{
thread_local std::string g_last_error;

template<typename F>
int guarded(F && f)
{
  try
  {
    f();
    return EXADG_OK;
  }
  catch(ExaDG::Error const & e)
  {
    g_last_error = e.what();
    return e.code;
  }
  catch(std::bad_alloc const &)
  {
    g_last_error = "host allocation failed";
    return EXADG_ERR_INTERNAL;
  }
  catch(std::exception const & e)
  {
    g_last_error = e.what();
    return EXADG_ERR_INTERNAL;
  }
  catch(...)
  {
    g_last_error = "unknown exception";
    return EXADG_ERR_INTERNAL;
  }
}

inline hipStream_t stream_of(void * s) { return static_cast<hipStream_t>(s); }

void require(bool cond, char const * msg)
{
  if(!cond)
    EXADG_THROW(EXADG_ERR_ARGUMENT, msg);
}
}  // namespace

struct exadg_laplace
{
  std::unique_ptr<ExaDG::Poisson::LaplaceOperatorBase> op;
  std::unique_ptr<ExaDG::Krylov::SolverCG>             cg;
};

struct exadg_guess
{
  std::unique_ptr<ExaDG::GuessNet> net;
};

struct exadg_divergence
{
  std::unique_ptr<ExaDG::DivergenceOperator> op;
  ExaDG::Reducer *                           red = nullptr;  // mean of singular right-hand sides
  ~exadg_divergence() { delete red; }
};

struct exadg_preconditioner
{
  int                                        kind = 0;
  int64_t                                    n    = 0;  // size of the operator it was built for
  std::unique_ptr<ExaDG::PreconditionerBase> p;
  ExaDG::DeviceVector<double>                work_r, work_c;  // exadg_preconditioner_smooth
};

namespace
{
ExaDG::Poisson::LaplaceOperator<double> & fp64_operator(exadg_laplace * op)
{
  auto * d = dynamic_cast<ExaDG::Poisson::LaplaceOperator<double> *>(op->op.get());
  if(!d)
    EXADG_THROW(EXADG_ERR_ARGUMENT, "the outer solver needs an FP64 operator");
  return *d;
}

void check_smooth_arguments(exadg_preconditioner const * p, int64_t n, double const * b,
                            double const * x)
{
  if(p->p == nullptr)
    EXADG_THROW(EXADG_ERR_ARGUMENT, "the preconditioner has no smoother");
  if(p->n != n)
    EXADG_THROW(EXADG_ERR_ARGUMENT, "the preconditioner was built for an operator of another size");
  if(b == x)
    EXADG_THROW(EXADG_ERR_ARGUMENT, "b and x must not alias");
}

// x += S (b - A x) with the residual in r and the correction in c (device, n_dofs each)
void smooth_polish(ExaDG::Poisson::LaplaceOperator<double> const & A, ExaDG::PreconditionerBase & P,
                   double const * b, double * x, double * r, double * c, hipStream_t s)
{
  A.residual(r, x, b, s);  // r = b - A x
  if(!P.smoother_vmult(c, r, s))
    EXADG_THROW(EXADG_ERR_ARGUMENT, "only the multigrid preconditioner has a smoother");
  ExaDG::vec::axpby<double>(x, 1.0, c, 1.0, A.m(), s);  // x += S r
}

// exadg_cg_solve and its no-allocation variant: the same CG loop, on the handle's own work
// vectors or on the static device workspace
int cg_solve_entry(exadg_laplace * op, exadg_preconditioner * prec, const double * b, double * x,
                   const exadg_solver_data * data, exadg_solve_info * info, void * stream,
                   bool noalloc)
{
  int                      status = EXADG_OK;
  ExaDG::Krylov::SolveInfo si;
  int const rc = guarded([&] {
    require(op && b && x && data, "null argument");
    require(data->max_iter >= 0 && data->rel_tol >= 0 && data->abs_tol >= 0, "invalid solver data");
    auto & A = fp64_operator(op);
    require(prec == nullptr || prec->n == A.m(),
            "the preconditioner was built for an operator of another size");
    ExaDG::Krylov::SolverData sd;
    sd.max_iter = data->max_iter;
    sd.abs_tol  = data->abs_tol;
    sd.rel_tol  = data->rel_tol;
    sd.rel_mode = data->rel_mode;
    hipStream_t const               s = stream_of(stream);
    ExaDG::PreconditionerBase * const P = prec ? prec->p.get() : nullptr;
    if(noalloc)
    {
      require(A.m() <= ExaDG::kNoAllocMaxDofs,
              "the operator has more DoFs than the no-allocation workspace holds "
              "(EXADG_NOALLOC_MAX_DOFS)");
      require(P == nullptr || !P->vmult_allocates(),
              "a multigrid preconditioner that replays HIP graphs captures them on first use; "
              "create it with capture_graph = 0 for the no-allocation solve");
      require(b != x, "b and x must not alias");
      int64_t const        N = ExaDG::kNoAllocMaxDofs;
      double * const       w = ExaDG::noalloc_workspace();
      ExaDG::Reducer       red(w + 4 * N);
      red.clear_counter(s);
      si = ExaDG::Krylov::cg_solve_in(A, P, b, x, sd, w, w + N, w + 2 * N, w + 3 * N, red, s);
    }
    else
    {
      if(!op->cg)
        op->cg = std::make_unique<ExaDG::Krylov::SolverCG>();
      si = op->cg->solve(A, P, b, x, sd, s);
    }
    if(si.not_finite)
    {
      g_last_error = "Last iteration step contained NaN or Inf values.";
      status       = EXADG_ERR_NOT_FINITE;
    }
    else if(!si.converged)
    {
      g_last_error = "CG reached the iteration limit";
      status       = EXADG_ERR_NO_CONVERGE;
    }
    else if(!si.true_residual_ok)
    {
      g_last_error = "CG converged on the recursive residual, but the true residual ||b - A x|| "
                     "exceeds 10x the tolerance";
      status       = EXADG_ERR_TRUE_RESIDUAL;
    }
  });
  if(info)
  {
    info->iterations = si.iterations;
    info->converged  = (si.converged && si.true_residual_ok && !si.not_finite) ? 1 : 0;
    info->not_finite = si.not_finite ? 1 : 0;
    info->r0         = si.r0;
    info->r_final    = si.r_final;
    info->b_norm     = si.b_norm;
    info->rho        = si.rho;
    info->n10        = si.n10;
    info->r_true     = si.r_true;
    info->tolerance  = si.tolerance;
  }
  return rc != EXADG_OK ? rc : status;
}

template<typename Number>
void mg_info(ExaDG::PreconditionerBase const * p, int * n_levels, double * lmax, int * degs,
             int * coarse_its)
{
  auto const & mg = static_cast<ExaDG::MultigridPreconditioner<Number> const &>(*p);
  if(n_levels)
    *n_levels = mg.n_levels();
  for(int l = 0; l < mg.n_levels(); ++l)
  {
    auto const & sm = l == 0 ? mg.coarse_chebyshev() : mg.smoother(l);
    if(lmax)
      lmax[l] = sm.lambda_max_estimate;
    if(degs)
      degs[l] = sm.degree;
  }
  if(coarse_its)
    *coarse_its = mg.last_coarse_iterations();
}

template<typename Number>
void mg_levels(ExaDG::PreconditionerBase const * p, int max_levels, int * n_levels, int * h,
               int * degree, int * dg, int64_t * n)
{
  auto const & mg = static_cast<ExaDG::MultigridPreconditioner<Number> const &>(*p);
  *n_levels       = mg.n_levels();
  for(int l = 0; l < mg.n_levels() && l < max_levels; ++l)
  {
    ExaDG::MGLevelInfo const i = mg.level_info(l);
    if(h)
      h[l] = i.h;
    if(degree)
      degree[l] = i.degree;
    if(dg)
      dg[l] = i.dg ? 1 : 0;
    if(n)
      n[l] = mg.level_operator(l).n();
  }
}
}  // namespace

extern "C" {

const char * exadg_last_error(void) { return g_last_error.c_str(); }

int exadg_laplace_create(const exadg_mesh_desc * mesh, const exadg_bc_desc * bc, int degree,
                         int precision, double ip_factor, exadg_laplace ** out)
{
  return guarded([&] {
    require(mesh && bc && out, "null argument");
    require(ip_factor > 0.0, "ip_factor must be positive");
    auto h = std::make_unique<exadg_laplace>();
    h->op  = ExaDG::Poisson::create_laplace_operator(*mesh, *bc, degree, precision, ip_factor);
    *out   = h.release();
  });
}

int exadg_laplace_destroy(exadg_laplace * op)
{
  return guarded([&] { delete op; });
}

int exadg_laplace_n_dofs(const exadg_laplace * op, int64_t * n_dofs)
{
  return guarded([&] {
    require(op && n_dofs, "null argument");
    *n_dofs = op->op->m();
  });
}

int exadg_laplace_info(const exadg_laplace * op, int64_t * n_cells, int * degree, int * precision,
                       int * is_affine)
{
  return guarded([&] {
    require(op, "null operator");
    if(n_cells)
      *n_cells = op->op->n_cells();
    if(degree)
      *degree = op->op->degree();
    if(precision)
      *precision = op->op->precision();
    if(is_affine)
      *is_affine = op->op->is_affine() ? 1 : 0;
  });
}

int exadg_laplace_vmult(exadg_laplace * op, void * dst, const void * src, void * stream)
{
  return guarded([&] {
    require(op, "null operator");
    op->op->vmult(dst, src, stream_of(stream));
  });
}

int exadg_laplace_vmult_add(exadg_laplace * op, void * dst, const void * src, void * stream)
{
  return guarded([&] {
    require(op, "null operator");
    op->op->vmult_add(dst, src, stream_of(stream));
  });
}

int exadg_laplace_calculate_inverse_diagonal(exadg_laplace * op, void * inv_diag, void * stream)
{
  return guarded([&] {
    require(op, "null operator");
    op->op->calculate_inverse_diagonal(inv_diag, stream_of(stream));
  });
}

int exadg_laplace_calculate_diagonal(exadg_laplace * op, void * diag, void * stream)
{
  return guarded([&] {
    require(op, "null operator");
    op->op->calculate_diagonal(diag, stream_of(stream));
  });
}

int exadg_laplace_rhs_sine(exadg_laplace * op, double omega, double * b, void * stream)
{
  return guarded([&] {
    require(op, "null operator");
    op->op->rhs_sine(omega, b, stream_of(stream));
  });
}

int exadg_laplace_rhs_n_points(const exadg_laplace * op, int64_t * n_cell_points,
                               int64_t * n_boundary_points)
{
  return guarded([&] {
    require(op && n_cell_points && n_boundary_points, "null argument");
    op->op->rhs_n_points(*n_cell_points, *n_boundary_points);
  });
}

int exadg_laplace_rhs_points(exadg_laplace * op, double * x_cell, double * x_boundary, void * stream)
{
  return guarded([&] {
    require(op, "null operator");
    op->op->rhs_points(x_cell, x_boundary, stream_of(stream));
  });
}

int exadg_laplace_rhs(exadg_laplace * op, const double * f, const double * boundary, double * b,
                      void * stream)
{
  return guarded([&] {
    require(op && b, "null argument");
    op->op->rhs(f, boundary, b, stream_of(stream));
  });
}

int exadg_mg_desc_default(int degree, exadg_mg_desc * out)
{
  return guarded([&] {
    require(out != nullptr, "null argument");
    require(degree >= 1 && degree <= 8, "degree must be in [1,8]");
    exadg_mg_desc d{};
    int           p = degree;
    while(true)
    {
      d.dg_degrees[d.n_dg_levels++] = p;
      if(p == 1)
        break;
      p = std::max(p / 2, 1);
    }
    d.coarse_q1            = 1;
    d.precision            = EXADG_FP32;
    d.smoother_degree      = 5;
    d.smoothing_range      = 20.0;
    d.eig_cg_iterations    = 20;
    d.coarse_solver        = EXADG_COARSE_CG;
    d.coarse_rel_tol       = 1e-3;
    d.coarse_max_iter      = 10000;
    d.operator_is_singular = 0;
    d.capture_graph        = 0;
    *out                   = d;
  });
}

int exadg_preconditioner_create(exadg_laplace * op, int kind, const exadg_mg_desc * mg,
                                exadg_preconditioner ** out, void * stream)
{
  return guarded([&] {
    require(op && out, "null argument");
    auto &     A = fp64_operator(op);
    auto       h = std::make_unique<exadg_preconditioner>();
    hipStream_t const s = stream_of(stream);
    h->kind             = kind;
    h->n                = A.m();
    if(kind == EXADG_PRECOND_NONE)
    {
    }
    else if(kind == EXADG_PRECOND_POINT_JACOBI)
    {
      h->p = std::make_unique<ExaDG::JacobiPreconditioner>(A);
      h->p->update(s);
    }
    else if(kind == EXADG_PRECOND_BLOCK_JACOBI || kind == EXADG_PRECOND_ADDITIVE_SCHWARZ)
    {
      // on the DG space the additive Schwarz preconditioner is the block-Jacobi inverse
      h->p = std::make_unique<ExaDG::BlockJacobiPreconditioner>(A);
      h->p->update(s);
    }
    else if(kind == EXADG_PRECOND_MULTIGRID)
    {
      require(mg != nullptr, "multigrid preconditioner needs an exadg_mg_desc");
      require(mg->n_dg_levels >= 1 && mg->n_dg_levels <= 16, "n_dg_levels out of range");
      require(mg->dg_degrees[0] == A.degree(), "dg_degrees[0] must equal the operator degree");
      ExaDG::MultigridData d;
      d.degrees.assign(mg->dg_degrees, mg->dg_degrees + mg->n_dg_levels);
      d.coarse_q1            = mg->coarse_q1 != 0;
      require(mg->mg_type >= EXADG_MG_LEGACY && mg->mg_type <= EXADG_MG_PHCMG, "unknown mg_type");
      require(mg->n_h_levels >= 0 && mg->n_h_levels <= 16, "n_h_levels out of range");
      d.type                 = static_cast<ExaDG::MultigridType>(mg->mg_type);
      d.n_h_levels           = mg->n_h_levels;
      require(mg->smoother_preconditioner >= EXADG_SMOOTHER_POINT_JACOBI &&
                mg->smoother_preconditioner <= EXADG_SMOOTHER_ADDITIVE_SCHWARZ,
              "unknown smoother_preconditioner");
      d.smoother_preconditioner =
        static_cast<ExaDG::SmootherPreconditioner>(mg->smoother_preconditioner);
      d.smoother_degree      = mg->smoother_degree;
      d.smoothing_range      = mg->smoothing_range;
      d.eig_cg_iterations    = mg->eig_cg_iterations;
      d.coarse_solver        = mg->coarse_solver == EXADG_COARSE_CHEBYSHEV ? ExaDG::CoarseSolver::chebyshev
                                                                           : ExaDG::CoarseSolver::cg;
      d.coarse_rel_tol       = mg->coarse_rel_tol;
      d.coarse_max_iter      = mg->coarse_max_iter;
      d.operator_is_singular = mg->operator_is_singular != 0;
      d.capture_graph        = mg->capture_graph != 0;
      require(d.smoother_degree >= 1 && d.smoothing_range > 1.0 && d.eig_cg_iterations >= 1,
              "invalid smoother data");
      require(d.coarse_rel_tol > 0.0 && d.coarse_rel_tol < 1.0 && d.coarse_max_iter >= 1,
              "invalid coarse solver data");
      h->p = ExaDG::create_multigrid(A.grid_ptr(), d, A.ip_factor(), mg->precision, s);
      h->p->update(s);
    }
    else
      EXADG_THROW(EXADG_ERR_ARGUMENT, "unknown preconditioner kind");
    h->kind = kind | (kind == EXADG_PRECOND_MULTIGRID && mg->precision == EXADG_FP32 ? 0x100 : 0);
    *out    = h.release();
  });
}

int exadg_preconditioner_vmult(exadg_preconditioner * p, double * dst, const double * src,
                               void * stream)
{
  return guarded([&] {
    require(p && dst && src, "null argument");
    require(dst != src, "dst and src must not alias");
    require(p->p != nullptr, "EXADG_PRECOND_NONE has no vmult");
    p->p->vmult(dst, src, stream_of(stream));
  });
}

int exadg_preconditioner_smooth(exadg_laplace * op, exadg_preconditioner * p, const double * b,
                                double * x, void * stream)
{
  return guarded([&] {
    require(op && p && b && x, "null argument");
    auto &        A = fp64_operator(op);
    int64_t const n = A.m();
    check_smooth_arguments(p, n, b, x);
    p->work_r.reinit(n);
    p->work_c.reinit(n);
    hipStream_t const s = stream_of(stream);
    smooth_polish(A, *p->p, b, x, p->work_r.data(), p->work_c.data(), s);
    HIP_CHECK(hipStreamSynchronize(s));
  });
}

int exadg_preconditioner_update(exadg_preconditioner * p, void * stream)
{
  return guarded([&] {
    require(p, "null argument");
    if(p->p)
      p->p->update(stream_of(stream));
  });
}

int exadg_preconditioner_destroy(exadg_preconditioner * p)
{
  return guarded([&] { delete p; });
}

int exadg_preconditioner_mg_info(const exadg_preconditioner * p, int * n_levels, double * lmax,
                                 int * cheb_degree, int * coarse_its)
{
  return guarded([&] {
    require(p, "null argument");
    require((p->kind & 0xff) == EXADG_PRECOND_MULTIGRID, "not a multigrid preconditioner");
    if(p->kind & 0x100)
      mg_info<float>(p->p.get(), n_levels, lmax, cheb_degree, coarse_its);
    else
      mg_info<double>(p->p.get(), n_levels, lmax, cheb_degree, coarse_its);
  });
}

int exadg_preconditioner_mg_levels(const exadg_preconditioner * p, int max_levels, int * n_levels,
                                   int * h, int * degree, int * dg, int64_t * n)
{
  return guarded([&] {
    require(p && n_levels, "null argument");
    require((p->kind & 0xff) == EXADG_PRECOND_MULTIGRID, "not a multigrid preconditioner");
    if(p->kind & 0x100)
      mg_levels<float>(p->p.get(), max_levels, n_levels, h, degree, dg, n);
    else
      mg_levels<double>(p->p.get(), max_levels, n_levels, h, degree, dg, n);
  });
}

int exadg_laplace_error_sine(exadg_laplace * op, double omega, const double * u, double * rel_error,
                             void * stream)
{
  return guarded([&] {
    require(op && u && rel_error, "null argument");
    *rel_error = op->op->error_sine(omega, u, stream_of(stream));
  });
}

int exadg_fourier_modes(uint64_t seed, int k_min, int k_max, int max_modes, double * modes,
                        int * n_modes)
{
  return guarded([&] {
    require(n_modes && k_min >= 1 && k_max >= k_min && k_max <= 16, "invalid Fourier shell");
    // the shell k_min <= |n| <= k_max, one vector of each pair +-n (first nonzero component
    // positive), in lexicographic order
    std::vector<std::array<int, 3>> shell;
    for(int a = -k_max; a <= k_max; ++a)
      for(int b = -k_max; b <= k_max; ++b)
        for(int c = -k_max; c <= k_max; ++c)
        {
          int const r2 = a * a + b * b + c * c;
          if(r2 < k_min * k_min || r2 > k_max * k_max)
            continue;
          if(a < 0 || (a == 0 && (b < 0 || (b == 0 && c < 0))))
            continue;
          shell.push_back({a, b, c});
        }
    *n_modes = (int)shell.size();
    if(modes == nullptr)
      return;
    require(max_modes >= (int)shell.size(), "mode buffer too small");
    uint64_t s    = seed;
    auto     next = [&s]() {
      uint64_t z = (s += 0x9E3779B97F4A7C15ull);
      z          = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z          = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      return (double)((z ^ (z >> 31)) >> 11) * (1.0 / 9007199254740992.0);
    };
    for(size_t m = 0; m < shell.size(); ++m)
    {
      int const *  n  = shell[m].data();
      double const kn = std::sqrt((double)(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]));
      // xi uniform on the sphere; d = (n/|n|) x xi, so div u = 0
      double const z = 2.0 * next() - 1.0, az = 2.0 * M_PI * next(), r = std::sqrt(1.0 - z * z);
      double const xi[3] = {r * std::cos(az), r * std::sin(az), z};
      double const kh[3] = {n[0] / kn, n[1] / kn, n[2] / kn};
      double *     md    = modes + 10 * m;
      md[0] = n[0], md[1] = n[1], md[2] = n[2];
      md[3] = kh[1] * xi[2] - kh[2] * xi[1];
      md[4] = kh[2] * xi[0] - kh[0] * xi[2];
      md[5] = kh[0] * xi[1] - kh[1] * xi[0];
      md[6] = std::pow(kn, -5.0 / 3.0);
      md[7] = 2.0 * M_PI * next();
      md[8] = kn;
      md[9] = 0.0;
    }
  });
}

int exadg_laplace_rhs_divergence_fourier(exadg_laplace * op, const double * modes, int n_modes,
                                         double t, double * b, void * stream)
{
  return guarded([&] {
    require(op, "null operator");
    op->op->rhs_divergence_fourier(modes, n_modes, t, b, stream_of(stream));
  });
}

int exadg_vector_random(void * v, int64_t n, int precision, uint64_t seed, void * stream)
{
  return guarded([&] {
    require(n >= 0, "negative length");
    if(precision != EXADG_FP64 && precision != EXADG_FP32)
      EXADG_THROW(EXADG_ERR_ARGUMENT, "precision must be EXADG_FP64 or EXADG_FP32");
    if(n == 0)
      return;
    require(v != nullptr, "null vector");
    if(precision == EXADG_FP64)
      ExaDG::vec::random_uniform<double>(static_cast<double *>(v), n, seed, stream_of(stream));
    else if(precision == EXADG_FP32)
      ExaDG::vec::random_uniform<float>(static_cast<float *>(v), n, seed, stream_of(stream));
    else
      EXADG_THROW(EXADG_ERR_ARGUMENT, "precision must be EXADG_FP64 or EXADG_FP32");
  });
}

int exadg_vector_subtract_mean(double * v, int64_t n, void * stream)
{
  return guarded([&] {
    require(v && n > 0, "invalid vector");
    ExaDG::Reducer red;
    ExaDG::vec::subtract_mean<double>(red, v, n, stream_of(stream));
    HIP_CHECK(hipStreamSynchronize(stream_of(stream)));
  });
}

int exadg_extrapolate(int order, const double * const * history, double * x0, int64_t n, void * stream)
{
  return guarded([&] {
    require(history && x0 && n > 0, "invalid argument");
    double beta[4];
    ExaDG::extrapolation_beta(order, beta);
    ExaDG::History<double> h;
    for(int k = 0; k < order; ++k)
    {
      require(history[k] != nullptr, "null history vector");
      h.p[k] = history[k];
    }
    hipError_t e = ExaDG::launch_combine(x0, h, order, beta, nullptr, 0.0, nullptr, n, nullptr,
                                         stream_of(stream));
    if(e != hipSuccess)
      EXADG_THROW(EXADG_ERR_HIP, hipGetErrorString(e));
  });
}

int exadg_divergence_create(const exadg_laplace * op, int degree_u, exadg_divergence ** out)
{
  return guarded([&] {
    require(op && out, "null argument");
    auto h = std::make_unique<exadg_divergence>();
    h->op  = std::make_unique<ExaDG::DivergenceOperator>(op->op->grid_ptr(), op->op->degree(),
                                                        degree_u);
    *out   = h.release();
  });
}

int exadg_divergence_destroy(exadg_divergence * div)
{
  return guarded([&] { delete div; });
}

int exadg_divergence_n_dofs(const exadg_divergence * div, int64_t * n_velocity, int64_t * n_pressure)
{
  return guarded([&] {
    require(div && n_velocity && n_pressure, "null argument");
    *n_velocity = div->op->n_velocity();
    *n_pressure = div->op->n_pressure();
  });
}

int exadg_divergence_apply(exadg_divergence * div, double * dst, const double * src, void * stream)
{
  return guarded([&] {
    require(div, "null operator");
    div->op->apply(dst, src, stream_of(stream));
  });
}

int exadg_divergence_velocity_points(exadg_divergence * div, double * x, void * stream)
{
  return guarded([&] {
    require(div, "null operator");
    div->op->velocity_points(x, stream_of(stream));
  });
}

int exadg_divergence_interpolate_fourier(exadg_divergence * div, const double * modes, int n_modes,
                                         double t, double * u, void * stream)
{
  return guarded([&] {
    require(div, "null operator");
    div->op->interpolate_fourier(modes, n_modes, t, u, stream_of(stream));
  });
}

int exadg_pressure_step(exadg_laplace * op, exadg_preconditioner * prec, exadg_divergence * div,
                        const double * velocity, double scale, int order,
                        const double * const * history, double * p_np, double * rhs,
                        const exadg_solver_data * data, exadg_solve_info * info, void * stream)
{
  int                      status = EXADG_OK;
  ExaDG::Krylov::SolveInfo si;
  int const rc = guarded([&] {
    require(op && div && velocity && p_np && rhs && data, "null argument");
    require(order >= 0 && order <= 4 && (order == 0 || history != nullptr), "invalid history");
    auto &            A = fp64_operator(op);
    hipStream_t const s = stream_of(stream);
    int64_t const     n = A.m();
    require(div->op->n_pressure() == n && div->op->degree_p() == A.degree(),
            "the divergence operator belongs to another pressure space");
    require(prec == nullptr || prec->n == n,
            "the preconditioner was built for an operator of another size");
    require(rhs != p_np && (void const *)rhs != (void const *)velocity, "rhs must not alias");
    // rhs_pressure (time_int_bdf_dual_splitting.cpp:536-545): rhs = scale * div(u_hat), with
    // scale = -gamma0 / dt; the pressure level of a singular problem is fixed by its mean
    div->op->apply(rhs, velocity, s);
    ExaDG::vec::axpby<double>(rhs, 0.0, rhs, scale, n, s);
    bool const singular = A.is_singular();
    if(!div->red)
      div->red = new ExaDG::Reducer();
    if(singular)
      ExaDG::vec::subtract_mean<double>(*div->red, rhs, n, s);
    // initial guess: extrapolation of the previous pressures (:487-501), or p_np as given
    if(order > 0)
    {
      double beta[4];
      ExaDG::extrapolation_beta(order, beta);
      ExaDG::History<double> h;
      for(int k = 0; k < order; ++k)
      {
        require(history[k] != nullptr && history[k] != p_np, "invalid history vector");
        h.p[k] = history[k];
      }
      hipError_t const e =
        ExaDG::launch_combine(p_np, h, order, beta, nullptr, 0.0, nullptr, n, nullptr, s);
      if(e != hipSuccess)
        EXADG_THROW(EXADG_ERR_HIP, hipGetErrorString(e));
    }
    if(!op->cg)
      op->cg = std::make_unique<ExaDG::Krylov::SolverCG>();
    ExaDG::Krylov::SolverData sd;
    sd.max_iter = data->max_iter;
    sd.abs_tol  = data->abs_tol;
    sd.rel_tol  = data->rel_tol;
    sd.rel_mode = data->rel_mode;
    si          = op->cg->solve(A, prec ? prec->p.get() : nullptr, rhs, p_np, sd, s);
    // adjust_pressure_level_if_undefined (:515-518): a singular problem keeps zero mean
    if(singular)
      ExaDG::vec::subtract_mean<double>(*div->red, p_np, n, s);
    HIP_CHECK(hipStreamSynchronize(s));
    if(si.not_finite)
    {
      g_last_error = "Last iteration step contained NaN or Inf values.";
      status       = EXADG_ERR_NOT_FINITE;
    }
    else if(!si.converged)
    {
      g_last_error = "CG reached the iteration limit";
      status       = EXADG_ERR_NO_CONVERGE;
    }
    else if(!si.true_residual_ok)
    {
      g_last_error = "CG converged on the recursive residual, but the true residual ||b - A x|| "
                     "exceeds 10x the tolerance";
      status       = EXADG_ERR_TRUE_RESIDUAL;
    }
  });
  if(info)
  {
    info->iterations = si.iterations;
    info->converged  = (si.converged && si.true_residual_ok && !si.not_finite) ? 1 : 0;
    info->not_finite = si.not_finite ? 1 : 0;
    info->r0         = si.r0;
    info->r_final    = si.r_final;
    info->b_norm     = si.b_norm;
    info->rho        = si.rho;
    info->n10        = si.n10;
    info->r_true     = si.r_true;
    info->tolerance  = si.tolerance;
  }
  return rc != EXADG_OK ? rc : status;
}

int exadg_guess_create(int degree, int history, int width, uint64_t seed, exadg_guess ** out)
{
  return guarded([&] {
    require(out != nullptr, "null argument");
    auto g = std::make_unique<exadg_guess>();
    g->net = std::make_unique<ExaDG::GuessNet>(degree, history, width, seed);
    *out   = g.release();
  });
}

int exadg_guess_destroy(exadg_guess * g)
{
  return guarded([&] { delete g; });
}

int exadg_guess_n_params(const exadg_guess * g, int * n_params, int * n_features, int * n_outputs)
{
  return guarded([&] {
    require(g, "null argument");
    if(n_params)
      *n_params = g->net->n_params();
    if(n_features)
      *n_features = g->net->n_features();
    if(n_outputs)
      *n_outputs = g->net->n_cell_dofs();
  });
}

int exadg_guess_get_params(const exadg_guess * g, float * dst, void * stream)
{
  return guarded([&] {
    require(g && dst, "null argument");
    hipStream_t const s = stream_of(stream);
    HIP_CHECK(hipMemcpyAsync(dst, g->net->params(), sizeof(float) * g->net->n_params(),
                             hipMemcpyDeviceToDevice, s));
    HIP_CHECK(hipStreamSynchronize(s));
  });
}

int exadg_guess_set_params(exadg_guess * g, const float * src, void * stream)
{
  return guarded([&] {
    require(g && src, "null argument");
    hipStream_t const s = stream_of(stream);
    HIP_CHECK(hipMemcpyAsync(g->net->params(), src, sizeof(float) * g->net->n_params(),
                             hipMemcpyDeviceToDevice, s));
    HIP_CHECK(hipStreamSynchronize(s));
  });
}

int exadg_guess_set_scales(exadg_guess * g, const double * scales, int n)
{
  return guarded([&] {
    require(g && scales && n == g->net->history() + 2, "expected history + 2 scales");
    g->net->set_scales(std::vector<double>(scales, scales + n));
  });
}

int exadg_guess_get_scales(const exadg_guess * g, double * scales, int n)
{
  return guarded([&] {
    require(g && scales && n == g->net->history() + 2, "expected history + 2 scales");
    std::vector<double> const v = g->net->scales();
    std::copy(v.begin(), v.end(), scales);
  });
}

int exadg_guess_loss_and_grad(exadg_guess * g, const float * X, const float * T, int B, float * grad,
                              double * loss, void * stream)
{
  return guarded([&] {
    require(g && X && T && B > 0, "invalid argument");
    hipStream_t const s = stream_of(stream);
    double const      l = g->net->loss_and_grad(X, T, B, s);
    if(loss)
      *loss = l;
    if(grad)
      HIP_CHECK(hipMemcpyAsync(grad, g->net->grads(), sizeof(float) * g->net->n_params(),
                               hipMemcpyDeviceToDevice, s));
  });
}

namespace
{
ExaDG::GuessLoss guess_loss(int kind)
{
  if(kind < EXADG_GUESS_LOSS_L2 || kind > EXADG_GUESS_LOSS_RESIDUAL)
    EXADG_THROW(EXADG_ERR_ARGUMENT, "unknown guess loss");
  return static_cast<ExaDG::GuessLoss>(kind);
}
}  // namespace

int exadg_guess_train(exadg_guess * g, exadg_laplace * op, int loss_kind, const double * P,
                      const double * Bv, int64_t n_dofs, int t_begin, int t_end, int epochs,
                      int batch_cells, double lr, uint64_t seed, double * loss, void * stream)
{
  return guarded([&] {
    require(g && P && Bv && n_dofs > 0 && epochs >= 0 && batch_cells > 0 && lr > 0,
            "invalid argument");
    ExaDG::GuessLoss const kind = guess_loss(loss_kind);
    require(kind == ExaDG::GuessLoss::l2 || op != nullptr, "this loss needs the operator");
    double const l = g->net->train(op ? op->op.get() : nullptr, kind, P, Bv, n_dofs, t_begin, t_end,
                                   epochs, batch_cells, (float)lr, seed, stream_of(stream));
    if(loss)
      *loss = l;
  });
}

int exadg_guess_evaluate(exadg_guess * g, exadg_laplace * op, int loss_kind, const double * P,
                         const double * Bv, int64_t n_dofs, int t_begin, int t_end, int batch_cells,
                         double * loss, void * stream)
{
  return guarded([&] {
    require(g && P && Bv && n_dofs > 0 && batch_cells > 0, "invalid argument");
    ExaDG::GuessLoss const kind = guess_loss(loss_kind);
    require(kind == ExaDG::GuessLoss::l2 || op != nullptr, "this loss needs the operator");
    double const l = g->net->evaluate(op ? op->op.get() : nullptr, kind, P, Bv, n_dofs, t_begin,
                                      t_end, batch_cells, stream_of(stream));
    if(loss)
      *loss = l;
  });
}

int exadg_guess_predict(exadg_guess * g, const double * const * history, const double * rhs,
                        double * x0, int64_t n_cells, int * used_fallback, void * stream)
{
  return guarded([&] {
    require(g && history && rhs && x0 && n_cells > 0, "invalid argument");
    ExaDG::History<double> h;
    for(int k = 0; k < g->net->history(); ++k)
      h.p[k] = history[k];
    int const bad = g->net->predict(x0, h, rhs, n_cells, 32768, stream_of(stream));
    if(used_fallback)
      *used_fallback = bad > 0 ? 1 : 0;
  });
}

int exadg_cg_solve(exadg_laplace * op, exadg_preconditioner * prec, const double * b, double * x,
                   const exadg_solver_data * data, exadg_solve_info * info, void * stream)
{
  return cg_solve_entry(op, prec, b, x, data, info, stream, false);
}

int exadg_cg_solve_noalloc(exadg_laplace * op, exadg_preconditioner * prec, const double * b,
                           double * x, const exadg_solver_data * data, exadg_solve_info * info,
                           void * stream)
{
  return cg_solve_entry(op, prec, b, x, data, info, stream, true);
}

}  // extern "C"

// ---------------------------------------------------------------- snapshots
#include "snapshot.h"

extern "C" {

int exadg_snapshot_write(const char * path, const exadg_snapshot_info * info,
                         const void * const * vectors, int on_device, int n_snapshots_keep,
                         void * stream)
{
  return guarded([&] {
    require(path && info, "null argument");
    ExaDG::write_snapshot(path, *info, vectors, on_device != 0, n_snapshots_keep, stream_of(stream));
  });
}

int exadg_snapshot_read_info(const char * path, exadg_snapshot_info * info)
{
  return guarded([&] {
    require(path && info, "null argument");
    *info = ExaDG::read_snapshot_info(path);
  });
}

int exadg_snapshot_read(const char * path, exadg_snapshot_info * info_out, void * const * vectors,
                        int n_vectors, int64_t n, int precision, int on_device, void * stream)
{
  return guarded([&] {
    require(path != nullptr, "null path");
    require(n_vectors >= 0 && n >= 0, "negative size");
    exadg_snapshot_info const info =
      ExaDG::read_snapshot(path, vectors, n_vectors, n, precision, on_device != 0, stream_of(stream));
    if(info_out)
      *info_out = info;
  });
}

}  // extern "C"
