// extern "C" boundary: every entry point catches, records the message and returns a status.
#include <cstring>
#include <memory>
#include <string>

#include "../../include/exadg_amd.h"
#include "common.h"
#include "laplace_operator.h"

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
};

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

}  // extern "C"
