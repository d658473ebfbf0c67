// A C++ consumer of the C ABI through include/exadg_amd.hpp, built by tests/test_cpp_consumer.py
// with g++ against libexadg_amd.so: the binding an ExaDG build would compile.
//   consumer cpu   -- calls that need no GPU (defaults, Fourier shell, error paths)
//   consumer gpu   -- a manufactured-solution solve on the device (CG + FP32 p-multigrid)
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "exadg_amd.hpp"

namespace
{
int fail(char const * what)
{
  std::printf("FAIL: %s\n", what);
  return 1;
}

exadg_mesh_desc box(int n)
{
  exadg_mesh_desc m{};
  for(int d = 0; d < 3; ++d)
  {
    m.n_cells[d] = n;
    m.lower[d]   = 0.0;
    m.upper[d]   = 1.0;
  }
  m.frequency      = 1;
  m.mapping_degree = 1;
  return m;
}

int run_cpu()
{
  exadg_mg_desc const mg = exadg_amd::default_multigrid_data(4);
  if(mg.n_dg_levels != 3 || mg.dg_degrees[0] != 4 || mg.dg_degrees[1] != 2 ||
     mg.dg_degrees[2] != 1 || mg.coarse_q1 != 1 || mg.precision != EXADG_FP32)
    return fail("multigrid defaults");
  int n_modes = 0;
  exadg_amd::check(exadg_fourier_modes(1, 1, 8, 0, nullptr, &n_modes));
  if(n_modes != 1054)
    return fail("Fourier shell size");
  exadg_bc_desc bc{};
  bool          thrown = false;
  try
  {
    exadg_amd::LaplaceOperator op(box(2), bc, 9);  // degree out of range
  }
  catch(exadg_amd::Error const & e)
  {
    thrown = e.code() == EXADG_ERR_ARGUMENT && std::strstr(e.what(), "degree") != nullptr;
  }
  if(!thrown)
    return fail("degree 9 must throw an argument error");
  if(exadg_vector_random(nullptr, 0, EXADG_FP64, 1, nullptr) != EXADG_OK)
    return fail("empty random vector");
  if(exadg_vector_random(nullptr, 4, 7, 1, nullptr) != EXADG_ERR_ARGUMENT)
    return fail("bad precision");
  std::printf("cpu ok\n");
  return 0;
}

int run_gpu()
{
  exadg_bc_desc const        bc{};  // homogeneous Dirichlet on all sides
  exadg_amd::LaplaceOperator op(box(4), bc, 3);
  int64_t const              n = op.n();
  double *                   b = nullptr, *x = nullptr;
  if(hipMalloc(&b, n * sizeof(double)) != hipSuccess || hipMalloc(&x, n * sizeof(double)) != hipSuccess)
    return fail("hipMalloc");
  if(hipMemset(x, 0, n * sizeof(double)) != hipSuccess)
    return fail("hipMemset");
  op.rhs_sine(M_PI, b);
  exadg_mg_desc const       mg = exadg_amd::default_multigrid_data(3);
  exadg_amd::Preconditioner prec(op, EXADG_PRECOND_MULTIGRID, &mg);
  exadg_solver_data const   data{1000, 1e-20, 1e-10, 1};
  exadg_solve_info          info{};
  unsigned int const        its = exadg_amd::solve(op, &prec, b, x, data, &info);
  double const              err = op.error_sine(M_PI, x);
  std::printf("gpu: n=%lld its=%u r_true/|b|=%.3e err=%.4e\n", (long long)n, its,
              info.r_true / info.b_norm, err);
  (void)hipFree(b);
  (void)hipFree(x);
  if(!(info.converged == 1 && info.r_true <= 1e-9 * info.b_norm && err < 1e-2 && its < 30))
    return fail("solve");
  std::printf("gpu ok\n");
  return 0;
}
}  // namespace

int main(int argc, char ** argv)
{
  try
  {
    if(argc > 1 && std::strcmp(argv[1], "gpu") == 0)
      return run_gpu();
    return run_cpu();
  }
  catch(exadg_amd::Error const & e)
  {
    std::printf("FAIL: exadg_amd::Error %d: %s\n", e.code(), e.what());
    return 1;
  }
}
