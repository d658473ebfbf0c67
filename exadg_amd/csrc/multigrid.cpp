#include "multigrid.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <string>

namespace ExaDG
{
// ---------------------------------------------------------------- helpers

void cg_tridiagonal_extreme_eigenvalues(std::vector<double> const & alpha,
                                        std::vector<double> const & beta, double & lmin,
                                        double & lmax)
{
  int const n = (int)alpha.size();
  if(n == 0)
  {
    lmin = lmax = 1.0;
    return;
  }
  std::vector<double> d(n), e(n, 0.0);
  for(int k = 0; k < n; ++k)
  {
    d[k] = 1.0 / alpha[k] + (k > 0 ? beta[k - 1] / alpha[k - 1] : 0.0);
    if(k + 1 < n)
      e[k] = std::sqrt(beta[k]) / alpha[k];
  }
  // Gershgorin bounds, then bisection with Sturm counts for the extreme eigenvalues
  double lo = d[0], hi = d[0];
  for(int k = 0; k < n; ++k)
  {
    double const r = (k > 0 ? std::fabs(e[k - 1]) : 0.0) + (k + 1 < n ? std::fabs(e[k]) : 0.0);
    lo             = std::min(lo, d[k] - r);
    hi             = std::max(hi, d[k] + r);
  }
  auto count_below = [&](double x) {
    int    c = 0;
    double q = 1.0;
    for(int k = 0; k < n; ++k)
    {
      q = d[k] - x - (k > 0 ? e[k - 1] * e[k - 1] / q : 0.0);
      if(q == 0.0)
        q = 1e-300;
      if(q < 0)
        ++c;
    }
    return c;
  };
  auto kth = [&](int k) {  // k-th smallest eigenvalue (0-based)
    double a = lo, b = hi;
    for(int it = 0; it < 200 && b - a > 1e-15 * std::max(1.0, std::fabs(b)); ++it)
    {
      double const m = 0.5 * (a + b);
      if(count_below(m) > k)
        b = m;
      else
        a = m;
    }
    return 0.5 * (a + b);
  };
  lmin = kth(0);
  lmax = kth(n - 1);
}

template<typename Number>
void JacobiData<Number>::apply(Number * z, Number const * r, int64_t n, hipStream_t s) const
{
  if(block == nullptr)
  {
    vec::scale(z, inv_diag, r, n, s);
    return;
  }
  BlockApplyArgs<Number> a;
  a.b   = r;
  a.dst = z;
  block->apply(a, s);
}

namespace
{
// host-synchronous PCG with point or block Jacobi on a level operator, recording the CG
// coefficients (setup-time eigenvalue estimates; coarse solves)
template<typename Number>
int level_pcg(LevelOperator<Number> const & op, JacobiData<Number> const & P, Number const * b, Number * x,
              double tol_abs, int max_iter, bool singular, Reducer & red, hipStream_t s,
              PcgWorkspace<Number> & ws, std::vector<double> * alphas = nullptr,
              std::vector<double> * betas = nullptr, bool x_is_zero = true)
{
  int64_t const n = op.n();
  ws.reinit(n);
  DeviceVector<Number> &r = ws.r, &z = ws.z, &p = ws.p, &v = ws.v;
  if(x_is_zero)
  {
    vec::set_zero(x, n, s);
    vec::copy(r.data(), b, n, s);
  }
  else
  {
    op.vmult(v.data(), x, s);
    vec::copy(r.data(), b, n, s);
    vec::axpby(r.data(), Number(-1), v.data(), Number(1), n, s);
  }
  if(singular)
    vec::subtract_mean(red, r.data(), n, s);
  P.apply(z.data(), r.data(), n, s);
  vec::norm2_and_dot(red, 1, 2, r.data(), z.data(), n, s);
  double const * h   = red.fetch(s, 4);
  double         rho = h[2];
  if(std::sqrt(h[1]) <= tol_abs)
    return 0;
  vec::copy(p.data(), z.data(), n, s);
  int it = 0;
  for(it = 1; it <= max_iter; ++it)
  {
    op.vmult(v.data(), p.data(), s);
    vec::dot(red, 0, p.data(), v.data(), n, s);
    double const pv    = red.fetch(s, 1)[0];
    double const alpha = rho / pv;
    vec::axpby(x, Number(alpha), p.data(), Number(1), n, s);
    vec::axpby(r.data(), Number(-alpha), v.data(), Number(1), n, s);
    if(singular)
      vec::subtract_mean(red, r.data(), n, s);
    P.apply(z.data(), r.data(), n, s);
    vec::norm2_and_dot(red, 1, 2, r.data(), z.data(), n, s);
    h                        = red.fetch(s, 4);
    double const rn          = std::sqrt(h[1]);
    double const rho_new     = h[2];
    double const beta        = rho_new / rho;
    if(alphas)
      alphas->push_back(alpha);
    if(betas)
      betas->push_back(beta);
    if(!std::isfinite(rn))
      EXADG_THROW(EXADG_ERR_NOT_FINITE, "non-finite residual in level CG");
    if(rn <= tol_abs)
      break;
    rho = rho_new;
    vec::axpby(p.data(), Number(1), z.data(), Number(beta), n, s);
  }
  return std::min(it, max_iter);
}

// DG level: the SIPG operator of one degree on the shared grid
template<typename Number>
class DGLevelOperator : public LevelOperator<Number>
{
public:
  DGLevelOperator(std::shared_ptr<StructuredGrid const> grid, int degree, double ip)
  {
    Poisson::LaplaceOperatorData d;
    d.degree    = degree;
    d.IP_factor = ip;
    op_         = std::make_unique<Poisson::LaplaceOperator<Number>>(std::move(grid), d);
  }
  int64_t n() const override { return op_->m(); }
  void    vmult(Number * dst, Number const * src, hipStream_t s) const override
  {
    op_->vmult(dst, src, s);
  }
  void inverse_diagonal(Number * d, hipStream_t s) const override
  {
    op_->calculate_inverse_diagonal(d, s);
  }
  std::unique_ptr<Poisson::CellBlockInverse<Number>> block_inverse(hipStream_t s) const override
  {
    return std::make_unique<Poisson::CellBlockInverse<Number>>(*op_, s);
  }
  void residual(Number * dst, Number const * src, Number const * b, hipStream_t s) const override
  {
    op_->residual(dst, src, b, s);
  }
  bool residual_restrict(Number * coarse, Number const * src, Number const * b, Number const * P,
                         int nc, hipStream_t s) const override
  {
    return op_->residual_restrict(coarse, src, b, P, nc, s);
  }
  bool chebyshev_step(Number * x_new, Number const * x, Number const * x_old, Number const * b,
                      Number const * dinv, Number f1, Number f2, bool has_old,
                      hipStream_t s) const override
  {
    op_->chebyshev_step(x_new, x, x_old, b, dinv, f1, f2, has_old, s);
    return true;
  }

private:
  std::unique_ptr<Poisson::LaplaceOperator<Number>> op_;
};
}  // namespace

template<typename Number>
void LevelOperator<Number>::residual(Number * dst, Number const * src, Number const * b,
                                     hipStream_t s) const
{
  if(dst == b)
    EXADG_THROW(EXADG_ERR_INTERNAL, "the unfused level residual needs dst != b");
  vmult(dst, src, s);
  vec::axpby(dst, Number(1), b, Number(-1), n(), s);  // dst = b - A src
}

// ---------------------------------------------------------------- Jacobi

JacobiPreconditioner::JacobiPreconditioner(Poisson::LaplaceOperator<double> const & op)
  : op_(op), inv_diag_(op.m())
{
}

void JacobiPreconditioner::update(hipStream_t s) { op_.calculate_inverse_diagonal(inv_diag_.data(), s); }

void JacobiPreconditioner::vmult(double * dst, double const * src, hipStream_t s)
{
  vec::scale(dst, inv_diag_.data(), src, inv_diag_.size(), s);
}

BlockJacobiPreconditioner::BlockJacobiPreconditioner(Poisson::LaplaceOperator<double> const & op)
  : op_(op)
{
}

void BlockJacobiPreconditioner::update(hipStream_t s)
{
  inv_ = std::make_unique<Poisson::CellBlockInverse<double>>(op_, s);
}

void BlockJacobiPreconditioner::vmult(double * dst, double const * src, hipStream_t s)
{
  if(dst == src)
    EXADG_THROW(EXADG_ERR_ARGUMENT, "dst and src must not alias");
  if(!inv_)
    EXADG_THROW(EXADG_ERR_NOT_READY, "the block-Jacobi preconditioner needs update()");
  BlockApplyArgs<double> a;
  a.b   = src;
  a.dst = dst;
  inv_->apply(a, s);
}

// ---------------------------------------------------------------- Q1 operator

template<typename Number>
Q1Operator<Number>::Q1Operator(StructuredGrid const & grid) : topo_(grid.topology())
{
  for(int d = 0; d < 3; ++d)
  {
    q_.nc[d]       = topo_.nc[d];
    q_.periodic[d] = topo_.periodic[d];
    q_.nn[d]       = topo_.nc[d] + (topo_.periodic[d] ? 0 : 1);
    for(int s = 0; s < 2; ++s)
      q_.dirichlet[2 * d + s] = (!topo_.periodic[d] && topo_.bc[2 * d + s] == EXADG_BC_DIRICHLET);
  }
  n_nodes_ = (int64_t)q_.nn[0] * q_.nn[1] * q_.nn[2];
  cg_      = make_cg_topology(topo_, 1);
  bool const shared = grid.affine();
  // element stiffness matrices int grad(phi_a) . grad(phi_b) with QGauss(2)
  std::vector<double> gp, gw;
  gauss_points(2, gp, gw);
  int64_t const nmat = shared ? 1 : grid.n_cells();
  mats_host_.assign((size_t)nmat * 64, 0.0);
  parallel_range(nmat, [&](int64_t cb, int64_t ce) {
    std::vector<double> x(3 * 8), J(9 * 8);
    for(int64_t c = cb; c < ce; ++c)
    {
      grid.evaluate(c, gp, gp, gp, x.data(), J.data());
      double * K = &mats_host_[(size_t)c * 64];
      for(int q = 0; q < 8; ++q)
      {
        double const xi[3] = {gp[q & 1], gp[(q >> 1) & 1], gp[(q >> 2) & 1]};
        double const w     = gw[q & 1] * gw[(q >> 1) & 1] * gw[(q >> 2) & 1];
        double       Jq[9];
        for(int k = 0; k < 9; ++k)
          Jq[k] = J[k * 8 + q];
        double const c00 = Jq[4] * Jq[8] - Jq[5] * Jq[7], c01 = Jq[5] * Jq[6] - Jq[3] * Jq[8],
                     c02 = Jq[3] * Jq[7] - Jq[4] * Jq[6];
        double const det = Jq[0] * c00 + Jq[1] * c01 + Jq[2] * c02;
        double       Ji[9];
        Ji[0] = c00 / det, Ji[3] = c01 / det, Ji[6] = c02 / det;
        Ji[1] = (Jq[2] * Jq[7] - Jq[1] * Jq[8]) / det;
        Ji[4] = (Jq[0] * Jq[8] - Jq[2] * Jq[6]) / det;
        Ji[7] = (Jq[1] * Jq[6] - Jq[0] * Jq[7]) / det;
        Ji[2] = (Jq[1] * Jq[5] - Jq[2] * Jq[4]) / det;
        Ji[5] = (Jq[2] * Jq[3] - Jq[0] * Jq[5]) / det;
        Ji[8] = (Jq[0] * Jq[4] - Jq[1] * Jq[3]) / det;
        double g[8][3];
        for(int a = 0; a < 8; ++a)
        {
          double gr[3];
          for(int e = 0; e < 3; ++e)
          {
            double v = 1.0;
            for(int d = 0; d < 3; ++d)
            {
              int const bit = (a >> d) & 1;
              if(d == e)
                v *= bit ? 1.0 : -1.0;
              else
                v *= bit ? xi[d] : 1.0 - xi[d];
            }
            gr[e] = v;
          }
          for(int i = 0; i < 3; ++i)  // J^{-T} grad_ref
            g[a][i] = Ji[0 * 3 + i] * gr[0] + Ji[1 * 3 + i] * gr[1] + Ji[2 * 3 + i] * gr[2];
        }
        for(int a = 0; a < 8; ++a)
          for(int b = 0; b < 8; ++b)
            K[a * 8 + b] += det * w * (g[a][0] * g[b][0] + g[a][1] * g[b][1] + g[a][2] * g[b][2]);
      }
    }
  });
  // assemble the 27-point stencil: node-centric gather over the (up to) 8 cells sharing the
  // node; row v of a cell's matrix couples its vertex v with the vertices m at offset
  // bits(m) - bits(v). Constrained neighbours are dropped (their columns read as zero),
  // constrained rows become identity rows.
  std::vector<char> const mask = constrained_mask();
  int64_t const           nn   = n_nodes_;
  std::vector<double>     W((size_t)27 * nn, 0.0), D((size_t)nn, 0.0);
  parallel_range(nn, [&](int64_t nb, int64_t ne) {
    for(int64_t node = nb; node < ne; ++node)
    {
      if(mask[node])
      {
        W[(size_t)13 * nn + node] = 1.0;
        D[node]                   = 1.0;
        continue;
      }
      int const nd[3] = {(int)(node % q_.nn[0]), (int)((node / q_.nn[0]) % q_.nn[1]),
                         (int)(node / ((int64_t)q_.nn[0] * q_.nn[1]))};
      for(int v = 0; v < 8; ++v)
      {
        int  cc[3];
        bool ok = true;
        for(int d = 0; d < 3; ++d)
        {
          int c = nd[d] - ((v >> d) & 1);
          if(c < 0)
            c = q_.periodic[d] ? c + q_.nc[d] : -1;
          if(c < 0 || c >= q_.nc[d])
            ok = false;
          cc[d] = c;
        }
        if(!ok)
          continue;
        int64_t const  cell = cell_id(topo_, cc);
        double const * K    = &mats_host_[(grid.affine() ? 0 : (size_t)cell * 64) + v * 8];
        D[node] += K[v];
        for(int m = 0; m < 8; ++m)
        {
          int nm[3], off[3];
          for(int d = 0; d < 3; ++d)
          {
            int const c = cc[d] + ((m >> d) & 1);
            nm[d]       = (q_.periodic[d] && c == q_.nc[d]) ? 0 : c;
            off[d]      = ((m >> d) & 1) - ((v >> d) & 1);
          }
          int64_t const id = nm[0] + (int64_t)q_.nn[0] * (nm[1] + (int64_t)q_.nn[1] * nm[2]);
          if(mask[id])
            continue;
          int const j = (off[0] + 1) + 3 * (off[1] + 1) + 9 * (off[2] + 1);
          W[(size_t)j * nn + node] += K[m];
        }
      }
    }
  });
  stencil_.upload(std::vector<Number>(W.begin(), W.end()));
  diag_.upload(std::vector<Number>(D.begin(), D.end()));
}

template<typename Number>
void Q1Operator<Number>::vmult(Number * dst, Number const * src, hipStream_t s) const
{
  hipError_t e = launch_q1_stencil<Number>(dst, src, stencil_.data(), q_, s);
  if(e != hipSuccess)
    EXADG_THROW(EXADG_ERR_HIP, hipGetErrorString(e));
}

template<typename Number>
bool Q1Operator<Number>::chebyshev_step(Number * x_new, Number const * x, Number const * x_old,
                                        Number const * b, Number const * dinv, Number f1, Number f2,
                                        bool has_old, hipStream_t s) const
{
  hipError_t e =
    launch_q1_chebyshev<Number>(x_new, x, x_old, b, dinv, stencil_.data(), q_, f1, f2, has_old, s);
  if(e != hipSuccess)
    EXADG_THROW(EXADG_ERR_HIP, hipGetErrorString(e));
  return true;
}

template<typename Number>
void Q1Operator<Number>::inverse_diagonal(Number * d, hipStream_t s) const
{
  vec::copy(d, diag_.data(), n_nodes_, s);
  Poisson::invert_diagonal_device<Number>(d, n_nodes_, s);
}

template<typename Number>
std::vector<char> Q1Operator<Number>::constrained_mask() const
{
  std::vector<char> m((size_t)n_nodes_, 0);
  for(int64_t i = 0; i < n_nodes_; ++i)
  {
    int nd[3] = {(int)(i % q_.nn[0]), (int)((i / q_.nn[0]) % q_.nn[1]), (int)(i / ((int64_t)q_.nn[0] * q_.nn[1]))};
    for(int d = 0; d < 3; ++d)
      if(!q_.periodic[d] && ((nd[d] == 0 && q_.dirichlet[2 * d]) || (nd[d] == q_.nc[d] && q_.dirichlet[2 * d + 1])))
        m[i] = 1;
  }
  return m;
}

// ---------------------------------------------------------------- continuous Q_p operator

template<typename Number>
CGOperator<Number>::CGOperator(StructuredGrid const & grid, int degree) : topo_(grid.topology())
{
  if(degree < 2 || degree > 8)
    EXADG_THROW(EXADG_ERR_ARGUMENT, "continuous levels of degree 2..8 (degree 1 is the Q1 level)");
  q_       = make_cg_topology(topo_, degree);
  n_nodes_ = (int64_t)q_.nn[0] * q_.nn[1] * q_.nn[2];
  n_cells_ = grid.n_cells();
  int const n = degree + 1;
  n3_         = n * n * n;
  ShapeInfo1D const sh = make_shape_info(degree);
  // S, Dq, w, then S.S, SD.SD, S.SD with SD = Dq S the nodal derivative at the Gauss points
  std::vector<double> SD((size_t)n * n, 0.0);
  for(int q = 0; q < n; ++q)
    for(int i = 0; i < n; ++i)
      for(int m = 0; m < n; ++m)
        SD[q * n + i] += sh.Dq[q * n + m] * sh.S[m * n + i];
  std::vector<Number> tab;
  for(double v : sh.S)
    tab.push_back((Number)v);
  for(double v : sh.Dq)
    tab.push_back((Number)v);
  for(double v : sh.qw)
    tab.push_back((Number)v);
  for(int e = 0; e < n * n; ++e)
    tab.push_back((Number)(sh.S[e] * sh.S[e]));
  for(int e = 0; e < n * n; ++e)
    tab.push_back((Number)(SD[e] * SD[e]));
  for(int e = 0; e < n * n; ++e)
    tab.push_back((Number)(sh.S[e] * SD[e]));
  tables_.upload(tab);
  double const vol = grid.h(0) * grid.h(1) * grid.h(2);
  for(int d = 0; d < 3; ++d)
    hscale_[d] = (Number)(vol / (grid.h(d) * grid.h(d)));
  if(!grid.affine())
  {
    GeometryData const g = compute_geometry(grid, sh, 1.0, true);
    metric_.upload(std::vector<Number>(g.cell_metric.begin(), g.cell_metric.end()));
  }
  cell_values_.reinit(n_cells_ * n3_);
}

template<typename Number>
CGCellArgs<Number> CGOperator<Number>::cell_args(Number const * src, bool diagonal) const
{
  CGCellArgs<Number> a;
  a.q      = q_;
  a.t      = topo_;
  a.tables = tables_.data();
  a.metric = metric_.size() ? metric_.data() : nullptr;
  for(int d = 0; d < 3; ++d)
    a.hscale[d] = hscale_[d];
  a.src     = diagonal ? nullptr : src;
  a.out     = cell_values_.data();
  a.n_cells = n_cells_;
  return a;
}

template<typename Number>
void CGOperator<Number>::assemble(CGAssembleArgs<Number> a, CGAssembleMode mode, hipStream_t s) const
{
  a.q           = q_;
  a.t           = topo_;
  a.cell_values = cell_values_.data();
  hipError_t e  = launch_cg_assemble<Number>(a, mode, s);
  if(e != hipSuccess)
    EXADG_THROW(EXADG_ERR_HIP, hipGetErrorString(e));
}

template<typename Number>
void CGOperator<Number>::vmult(Number * dst, Number const * src, hipStream_t s) const
{
  hipError_t e = launch_cg_cell<Number>(cell_args(src, false), false, s);
  if(e != hipSuccess)
    EXADG_THROW(EXADG_ERR_HIP, hipGetErrorString(e));
  CGAssembleArgs<Number> a;
  a.dst = dst;
  a.x   = src;
  assemble(a, CGAssembleMode::vmult, s);
}

template<typename Number>
void CGOperator<Number>::residual(Number * dst, Number const * src, Number const * b,
                                  hipStream_t s) const
{
  hipError_t e = launch_cg_cell<Number>(cell_args(src, false), false, s);
  if(e != hipSuccess)
    EXADG_THROW(EXADG_ERR_HIP, hipGetErrorString(e));
  CGAssembleArgs<Number> a;
  a.dst = dst;
  a.x   = src;
  a.b   = b;
  assemble(a, CGAssembleMode::residual, s);
}

template<typename Number>
bool CGOperator<Number>::chebyshev_step(Number * x_new, Number const * x, Number const * x_old,
                                        Number const * b, Number const * dinv, Number f1, Number f2,
                                        bool has_old, hipStream_t s) const
{
  hipError_t e = launch_cg_cell<Number>(cell_args(x, false), false, s);
  if(e != hipSuccess)
    EXADG_THROW(EXADG_ERR_HIP, hipGetErrorString(e));
  CGAssembleArgs<Number> a;
  a.dst     = x_new;
  a.x       = x;
  a.b       = b;
  a.x_old   = x_old;
  a.dinv    = dinv;
  a.f1      = f1;
  a.f2      = f2;
  a.has_old = has_old ? 1 : 0;
  assemble(a, CGAssembleMode::chebyshev, s);
  return true;
}

template<typename Number>
void CGOperator<Number>::inverse_diagonal(Number * d, hipStream_t s) const
{
  hipError_t e = launch_cg_cell<Number>(cell_args(nullptr, true), true, s);
  if(e != hipSuccess)
    EXADG_THROW(EXADG_ERR_HIP, hipGetErrorString(e));
  CGAssembleArgs<Number> a;
  a.dst = d;
  assemble(a, CGAssembleMode::diagonal, s);
  Poisson::invert_diagonal_device<Number>(d, n_nodes_, s);
}

// ---------------------------------------------------------------- Chebyshev

template<typename Number>
JacobiData<Number> ChebyshevSmoother<Number>::jacobi_data() const
{
  JacobiData<Number> P;
  P.inv_diag = inv_diag_.data();
  P.block    = block_.get();
  return P;
}

template<typename Number>
void ChebyshevSmoother<Number>::setup(LevelOperator<Number> const & op, int degree_,
                                      double smoothing_range, int eig_cg_iterations, hipStream_t s,
                                      SmootherPreconditioner prec)
{
  op_          = &op;
  degree       = degree_;
  int64_t const n = op.n();
  t_.reinit(n);
  block_.reset();
  if(prec == SmootherPreconditioner::point_jacobi)
  {
    inv_diag_.reinit(n);
    op.inverse_diagonal(inv_diag_.data(), s);
  }
  else
  {
    // PreconditionerSmoother::BlockJacobi; AdditiveSchwarz is the same inverse on DG levels
    if(op.continuous())
      EXADG_THROW(EXADG_ERR_ARGUMENT,
                  prec == SmootherPreconditioner::block_jacobi
                    ? "Block Jacobi only implemented for DG!"
                    : "additive Schwarz on continuous levels is not implemented");
    block_ = op.block_inverse(s);
  }
  // start vector of deal.II's eigenvalue estimate: (i mod 11) minus its mean (the smoother
  // carries no constraints object, so constrained entries are not zeroed);
  // eig_cg_iterations steps of Jacobi-PCG, Lanczos tridiagonal
  std::vector<Number> h((size_t)n);
  double              mean = 0;
  for(int64_t i = 0; i < n; ++i)
    mean += (double)(i % 11);
  mean /= (double)n;
  double nrm2 = 0;
  for(int64_t i = 0; i < n; ++i)
  {
    double const v = (double)(i % 11) - mean;
    h[i]           = (Number)v;
    nrm2 += v * v;
  }
  DeviceVector<Number> b(n), x(n);
  HIP_CHECK(hipMemcpyAsync(b.data(), h.data(), n * sizeof(Number), hipMemcpyHostToDevice, s));
  Reducer              red;
  PcgWorkspace<Number> ws;
  std::vector<double>  al, be;
  cg_iterations = level_pcg<Number>(op, jacobi_data(), b.data(), x.data(), 1e-2 * std::sqrt(nrm2),
                                    eig_cg_iterations, false, red, s, ws, &al, &be);
  cg_tridiagonal_extreme_eigenvalues(al, be, lambda_min_estimate, lambda_max_estimate);
  // interval [lambda_max/range, lambda_max] with a 1.2 safety factor on the CG estimate
  double const max_ev = 1.2 * lambda_max_estimate;
  double const alpha  = max_ev / smoothing_range;
  delta               = 0.5 * (max_ev - alpha);
  theta               = 0.5 * (max_ev + alpha);
}

template<typename Number>
void ChebyshevSmoother<Number>::setup_coarse(LevelOperator<Number> const & op, double rel_tol,
                                             bool singular, hipStream_t s)
{
  // MGCoarseChebyshev::update (coarse_grid_solvers.h:249-301) with estimate_eigenvalues_cg
  // (utilities/compute_eigenvalues.h:38-88): srand(1) right-hand side, CG to 1e-5
  op_             = &op;
  int64_t const n = op.n();
  inv_diag_.reinit(n);
  t_.reinit(n);
  op.inverse_diagonal(inv_diag_.data(), s);
  std::vector<Number> h((size_t)n);
  srand(1);
  double mean = 0;
  std::vector<double> hv((size_t)n);
  for(int64_t i = 0; i < n; ++i)
  {
    hv[i] = (double)rand() / RAND_MAX;
    mean += hv[i];
  }
  mean /= (double)n;
  double nrm2 = 0;
  for(int64_t i = 0; i < n; ++i)
  {
    double const v = singular ? hv[i] - mean : hv[i];
    h[i]           = (Number)v;
    nrm2 += v * v;
  }
  DeviceVector<Number> b(n), x(n);
  HIP_CHECK(hipMemcpyAsync(b.data(), h.data(), n * sizeof(Number), hipMemcpyHostToDevice, s));
  Reducer              red;
  PcgWorkspace<Number> ws;
  std::vector<double>  al, be;
  cg_iterations = level_pcg<Number>(op, jacobi_data(), b.data(), x.data(), 1e-5 * std::sqrt(nrm2),
                                    10000, singular, red, s, ws, &al, &be);
  cg_tridiagonal_extreme_eigenvalues(al, be, lambda_min_estimate, lambda_max_estimate);
  double const factor = 1.1;
  double const max_ev = factor * lambda_max_estimate;
  double const range  = lambda_max_estimate / lambda_min_estimate * factor;
  double const sigma  = (1. - std::sqrt(1. / range)) / (1. + std::sqrt(1. / range));
  double const eps    = rel_tol;
  degree = (int)(std::log(1. / eps + std::sqrt(1. / eps / eps - 1.)) / std::log(1. / sigma));
  degree = std::max(degree, 1);
  double const alpha = max_ev / range;
  delta              = 0.5 * (max_ev - alpha);
  theta              = 0.5 * (max_ev + alpha);
}

template<typename Number>
void ChebyshevSmoother<Number>::vmult(Number *& x, Number *& x_alt, Number const * b,
                                      hipStream_t s) const
{
  int64_t const n = op_->n();
  if(block_)
  {
    BlockApplyArgs<Number> a;
    a.b   = b;
    a.dst = x;
    a.f2  = Number(1. / theta);
    block_->apply(a, s);
  }
  else
    vec::chebyshev_update<Number>(x, nullptr, nullptr, b, nullptr, inv_diag_.data(), Number(0),
                                  Number(1. / theta), true, false, n, s);
  if(degree < 2 || std::fabs(delta) < 1e-40)
    return;
  double rhok = delta / theta, sigma = theta / delta;
  for(int i = 1; i < degree; ++i)
  {
    double const rhokp = 1. / (2. * sigma - rhok);
    double const f1 = rhokp * rhok, f2 = 2. * rhokp / delta;
    rhok            = rhokp;
    apply_step(x_alt, x, x_alt, b, Number(f1), Number(f2), i > 1, s);
    std::swap(x, x_alt);
  }
}

template<typename Number>
void ChebyshevSmoother<Number>::step(Number *& x, Number *& x_alt, Number const * b,
                                     hipStream_t s) const
{
  apply_step(x_alt, x, x_alt, b, Number(0), Number(1. / theta), false, s);
  std::swap(x, x_alt);
  if(degree < 2 || std::fabs(delta) < 1e-40)
    return;
  double rhok = delta / theta, sigma = theta / delta;
  for(int i = 1; i < degree; ++i)
  {
    double const rhokp = 1. / (2. * sigma - rhok);
    double const f1 = rhokp * rhok, f2 = 2. * rhokp / delta;
    rhok            = rhokp;
    apply_step(x_alt, x, x_alt, b, Number(f1), Number(f2), true, s);
    std::swap(x, x_alt);
  }
}

template<typename Number>
void ChebyshevSmoother<Number>::apply_step(Number * x_new, Number const * x, Number const * x_old,
                                           Number const * b, Number f1, Number f2, bool has_old,
                                           hipStream_t s) const
{
  if(block_)
  {
    // t = A x, then the block inverse with the Chebyshev update in its epilogue
    op_->vmult(t_.data(), x, s);
    BlockApplyArgs<Number> a;
    a.b       = b;
    a.t       = t_.data();
    a.x       = x;
    a.x_old   = x_old;
    a.dst     = x_new;
    a.f1      = f1;
    a.f2      = f2;
    a.variant = has_old ? 2 : 1;
    block_->apply(a, s);
    return;
  }
  if(op_->chebyshev_step(x_new, x, x_old, b, inv_diag_.data(), f1, f2, has_old, s))
    return;
  op_->vmult(t_.data(), x, s);
  vec::chebyshev_update<Number>(x_new, x, x_old, b, t_.data(), inv_diag_.data(), f1, f2, false,
                                has_old, op_->n(), s);
}

// ---------------------------------------------------------------- multigrid

std::vector<MGLevelInfo> multigrid_levels(MultigridType type, std::vector<int> const & degrees,
                                          int n_h_levels, bool coarse_q1)
{
  using T = MultigridType;
  if(degrees.empty())
    EXADG_THROW(EXADG_ERR_ARGUMENT, "multigrid needs a p-sequence");
  for(size_t i = 1; i < degrees.size(); ++i)
    if(degrees[i] >= degrees[i - 1] || degrees[i] < 1)
      EXADG_THROW(EXADG_ERR_ARGUMENT, "degrees must decrease strictly from fine to coarse");
  if(type == T::legacy)
  {
    if(coarse_q1 && degrees.back() != 1)
      EXADG_THROW(EXADG_ERR_ARGUMENT, "the continuous Q1 level must follow a DG degree-1 level");
    type = coarse_q1 ? T::pcMG : T::pMG;
  }
  // which transfers the type involves, and where the c-transfer sits relative to p
  char const * const name_of[] = {"", "h", "ch", "hc", "c", "p", "cp", "pc", "hp", "chp", "hcp",
                                  "hpc", "ph", "cph", "pch", "phc"};
  int const ti = (int)type;
  if(ti < 1 || ti > 15)
    EXADG_THROW(EXADG_ERR_ARGUMENT, "unknown multigrid type");
  std::string const name = name_of[ti];
  bool const has_h = name.find('h') != std::string::npos;
  bool const has_p = name.find('p') != std::string::npos;
  bool const has_c = name.find('c') != std::string::npos;
  int const  H     = has_h ? std::max(1, n_h_levels) : 1;
  int const  k     = degrees[0];
  // p-levels coarse to fine (the mesh is filled in below)
  std::vector<MGLevelInfo> P;
  if(!has_p)
  {
    if(has_c)
      P.push_back({0, k, false});
    P.push_back({0, k, true});
  }
  else
  {
    bool const c_first = has_c && name.find('c') < name.find('p');
    bool const c_last  = has_c && !c_first;
    if(c_last)
      P.push_back({0, degrees.back(), false});
    for(auto it = degrees.rbegin(); it != degrees.rend(); ++it)
      P.push_back({0, *it, !c_first});
    if(c_first)
      P.push_back({0, k, true});
  }
  int const                np = (int)P.size();
  std::vector<MGLevelInfo> L;
  auto on = [&](int h, MGLevelInfo p) { L.push_back({h, p.degree, p.dg}); };
  switch(type)
  {
    case T::hMG:
      for(int h = 0; h < H; ++h)
        on(h, P[0]);
      break;
    case T::cMG:
    case T::pMG:
    case T::cpMG:
    case T::pcMG:
      for(auto const & p : P)
        on(H - 1, p);
      break;
    case T::chMG:
      for(int h = 0; h < H; ++h)
        on(h, P[0]);
      on(H - 1, P[1]);
      break;
    case T::hcMG:
      on(0, P[0]);
      for(int h = 0; h < H; ++h)
        on(h, P[1]);
      break;
    case T::phMG:
    case T::cphMG:
    case T::pchMG:
      for(int h = 0; h + 1 < H; ++h)
        on(h, P[0]);
      for(auto const & p : P)
        on(H - 1, p);
      break;
    case T::hpMG:
    case T::hcpMG:
    case T::hpcMG:
      for(int i = 0; i + 1 < np; ++i)
        on(0, P[i]);
      for(int h = 0; h < H; ++h)
        on(h, P[np - 1]);
      break;
    case T::phcMG:
      on(0, P[0]);
      for(int h = 0; h + 1 < H; ++h)
        on(h, P[1]);
      for(int i = 1; i < np; ++i)
        on(H - 1, P[i]);
      break;
    case T::chpMG:
      for(int i = 0; i + 2 < np; ++i)
        on(0, P[i]);
      for(int h = 0; h < H; ++h)
        on(h, P[np - 2]);
      on(H - 1, P[np - 1]);
      break;
    default:
      EXADG_THROW(EXADG_ERR_ARGUMENT, "unknown multigrid type");
  }
  for(size_t l = 1; l < L.size(); ++l)
    if((L[l].h != L[l - 1].h) + (L[l].degree != L[l - 1].degree) + (L[l].dg != L[l - 1].dg) != 1)
      EXADG_THROW(EXADG_ERR_ARGUMENT,
                  "between two multigrid levels exactly one of mesh, degree, continuity may change");
  return L;
}

namespace
{
// 1D continuous transfer from (nc_c cells, degree pc) to (nc_f cells, degree pf) on GL nodes,
// nc_f = nc_c (p) or 2 nc_c (h): row I of the prolongation interpolates the coarse cell that
// contains fine node I; entries of magnitude below 1e-14 (coinciding nodes) are dropped
struct Line1D
{
  int                 rows = 0, width = 0;
  std::vector<int>    idx;
  std::vector<double> w;
};

Line1D line_prolongation(int nc_c, int pc, int nc_f, int pf, bool periodic)
{
  Line1D L;
  int const rows_c = nc_c * pc + (periodic ? 0 : 1);
  L.rows           = nc_f * pf + (periodic ? 0 : 1);
  L.width          = pc + 1;
  std::vector<double> xf, xc;
  gauss_lobatto_points(pf + 1, xf);
  gauss_lobatto_points(pc + 1, xc);
  std::vector<double> val(pc + 1);
  L.idx.assign((size_t)L.rows * L.width, 0);
  L.w.assign((size_t)L.rows * L.width, 0.0);
  for(int I = 0; I < L.rows; ++I)
  {
    int const    cf = std::min(I / pf, nc_f - 1), a = I - cf * pf;
    int const    parent = nc_f == nc_c ? cf : cf / 2;
    double const xi     = nc_f == nc_c ? xf[a] : 0.5 * ((cf % 2) + xf[a]);
    lagrange_basis(xc, xi, val.data(), nullptr);
    for(int j = 0; j <= pc; ++j)
    {
      int const c                  = parent * pc + j;
      L.idx[(size_t)I * L.width + j] = periodic ? c % rows_c : c;
      L.w[(size_t)I * L.width + j]   = std::fabs(val[j]) < 1e-14 ? 0.0 : val[j];
    }
  }
  return L;
}

Line1D line_transpose(Line1D const & P, int rows_c)
{
  std::vector<std::vector<std::pair<int, double>>> col((size_t)rows_c);
  for(int I = 0; I < P.rows; ++I)
    for(int j = 0; j < P.width; ++j)
    {
      double const w = P.w[(size_t)I * P.width + j];
      if(w != 0.0)
        col[P.idx[(size_t)I * P.width + j]].push_back({I, w});
    }
  Line1D R;
  R.rows  = rows_c;
  R.width = 1;
  for(auto const & c : col)
    R.width = std::max(R.width, (int)c.size());
  R.idx.assign((size_t)R.rows * R.width, 0);
  R.w.assign((size_t)R.rows * R.width, 0.0);
  for(int a = 0; a < rows_c; ++a)
    for(size_t j = 0; j < col[a].size(); ++j)
    {
      R.idx[(size_t)a * R.width + j] = col[a][j].first;
      R.w[(size_t)a * R.width + j]   = col[a][j].second;
    }
  return R;
}

int automatic_h_levels(StructuredGrid const & g)
{
  int H = 1;
  for(;;)
  {
    bool ok = true;
    for(int d = 0; d < 3; ++d)
    {
      int const n = g.mesh().n_cells[d];
      if(n % (1 << H) != 0 || (n >> H) < 2)
        ok = false;
    }
    if(!ok)
      return H;
    ++H;
  }
}
}  // namespace

template<typename Number>
MultigridPreconditioner<Number>::MultigridPreconditioner(std::shared_ptr<StructuredGrid const> grid,
                                                         MultigridData const & data, double ip,
                                                         hipStream_t s)
  : grid_(std::move(grid)), data_(data)
{
  int const n_h = data.n_h_levels > 0 ? data.n_h_levels : automatic_h_levels(*grid_);
  std::vector<MGLevelInfo> const info =
    multigrid_levels(data.type, data.degrees, n_h, data.coarse_q1);
  int const n_levels = (int)info.size();
  int const H        = info.back().h + 1;
  std::vector<std::shared_ptr<StructuredGrid const>> grids((size_t)H);
  grids[H - 1] = grid_;
  for(int h = 0; h + 1 < H; ++h)
    grids[h] = std::make_shared<StructuredGrid const>(grid_->coarsened(H - 1 - h));
  levels_.resize(n_levels);
  for(int l = 0; l < n_levels; ++l)
  {
    Level & L = levels_[l];
    L.h       = info[l].h;
    L.degree  = info[l].degree;
    L.dg      = info[l].dg;
    L.grid    = grids[L.h];
    if(L.dg)
      L.op = std::make_unique<DGLevelOperator<Number>>(L.grid, L.degree, ip);
    else if(L.degree == 1)
      L.op = std::make_unique<Q1Operator<Number>>(*L.grid);
    else
      L.op = std::make_unique<CGOperator<Number>>(*L.grid, L.degree);
  }
  for(int k = 0; k < n_levels; ++k)
  {
    Level &       L = levels_[k];
    int64_t const n = L.op->n();
    L.defect.reinit(n);
    L.sol_a.reinit(n);
    L.sol_b.reinit(n);
    L.t.reinit(n);
    L.sol     = L.sol_a.data();
    L.sol_alt = L.sol_b.data();
    vec::set_zero(L.sol, n, s);
    vec::set_zero(L.sol_alt, n, s);
    if(k == 0)
      continue;
    Level const & C = levels_[k - 1];
    if(C.dg != L.dg)
    {
      if(C.dg)
        EXADG_THROW(EXADG_ERR_ARGUMENT, "a c-transfer needs the continuous level below the DG one");
      L.transfer = 'c';
      continue;
    }
    L.transfer = C.degree != L.degree ? 'p' : 'h';
    if(L.dg && L.transfer == 'p')
    {
      // P[a][i] = phi^coarse_i(x^fine_a) on Gauss-Lobatto nodes
      int const           nf = L.degree + 1, nc = C.degree + 1;
      std::vector<double> xf, xc;
      gauss_lobatto_points(nf, xf);
      gauss_lobatto_points(nc, xc);
      std::vector<Number> P((size_t)nf * nc);
      std::vector<double> val(nc);
      for(int a = 0; a < nf; ++a)
      {
        lagrange_basis(xc, xf[a], val.data(), nullptr);
        for(int i = 0; i < nc; ++i)
          P[a * nc + i] = (Number)val[i];
      }
      L.P.upload(P);
    }
    else if(L.dg)
    {
      // child b of a parent: Ph[b][a][i] = phi_i((b + x_a) / 2)
      int const           n = L.degree + 1;
      std::vector<double> x, val(n);
      gauss_lobatto_points(n, x);
      std::vector<Number> P((size_t)2 * n * n);
      for(int b = 0; b < 2; ++b)
        for(int a = 0; a < n; ++a)
        {
          lagrange_basis(x, 0.5 * (b + x[a]), val.data(), nullptr);
          for(int i = 0; i < n; ++i)
            P[(size_t)b * n * n + a * n + i] = (Number)val[i];
        }
      L.P.upload(P);
    }
    else
    {
      CGTopology const & qf = *L.op->cg_topology();
      CGTopology const & qc = *C.op->cg_topology();
      std::vector<int>    idx;
      std::vector<double> w;
      Line1D              lines[6];
      for(int d = 0; d < 3; ++d)
      {
        lines[d] = line_prolongation(qc.nc[d], qc.p, qf.nc[d], qf.p, qf.periodic[d] != 0);
        lines[3 + d] = line_transpose(lines[d], qc.nn[d]);
      }
      std::vector<size_t> off(6);
      for(int j = 0; j < 6; ++j)
      {
        off[j] = idx.size();
        idx.insert(idx.end(), lines[j].idx.begin(), lines[j].idx.end());
        w.insert(w.end(), lines[j].w.begin(), lines[j].w.end());
      }
      L.line_idx.upload(idx);
      L.line_w.upload(std::vector<Number>(w.begin(), w.end()));
      for(int j = 0; j < 6; ++j)
      {
        LineOperator<Number> & o = j < 3 ? L.line_prolong[j] : L.line_restrict[j - 3];
        o.rows                   = lines[j].rows;
        o.width                  = lines[j].width;
        o.idx                    = L.line_idx.data() + off[j];
        o.w                      = L.line_w.data() + off[j];
      }
      int64_t const f0 = qf.nn[0], f1 = qf.nn[1], f2 = qf.nn[2];
      int64_t const c0 = qc.nn[0], c1 = qc.nn[1], c2 = qc.nn[2];
      L.work_a.reinit(std::max(f0 * c1 * c2, c0 * f1 * f2));
      L.work_b.reinit(std::max(f0 * f1 * c2, c0 * c1 * f2));
    }
  }
  int64_t const n0 = levels_[0].op->n();
  coarse_inv_diag_.reinit(n0);
  cr_.reinit(n0);
  coarse_ws_.reinit(n0);
  // PreconditionerBase::update() performs the smoother and coarse-solver setup
}

template<typename Number>
MultigridPreconditioner<Number>::~MultigridPreconditioner()
{
  clear_graphs();
  if(capture_stream_)
    (void)hipStreamDestroy(capture_stream_);
}

template<typename Number>
void MultigridPreconditioner<Number>::clear_graphs()
{
  for(Graph & g : graphs_)
    (void)hipGraphExecDestroy(g.exec);
  graphs_.clear();
  eager_calls_ = 0;
}

template<typename Number>
void MultigridPreconditioner<Number>::update(hipStream_t s)
{
  clear_graphs();  // the captured kernels hold the Chebyshev coefficients as arguments
  for(int l = 1; l < n_levels(); ++l)
    levels_[l].smoother.setup(*levels_[l].op, data_.smoother_degree, data_.smoothing_range,
                              data_.eig_cg_iterations, s, data_.smoother_preconditioner);
  if(data_.coarse_solver == CoarseSolver::chebyshev)
    coarse_cheb_.setup_coarse(*levels_[0].op, data_.coarse_rel_tol, data_.operator_is_singular, s);
  else
    levels_[0].op->inverse_diagonal(coarse_inv_diag_.data(), s);
  HIP_CHECK(hipStreamSynchronize(s));
}

template<typename Number>
void MultigridPreconditioner<Number>::prolongate_and_add(int l, Number * dst, Number const * src,
                                                         hipStream_t s) const
{
  Level const & F = levels_[l];
  Level const & C = levels_[l - 1];
  hipError_t    e = hipSuccess;
  if(F.transfer == 'c')
    e = launch_c_prolongate_add<Number>(dst, src, F.grid->topology(), *C.op->cg_topology(), s);
  else if(F.dg && F.transfer == 'p')
    e = launch_p_prolongate_add<Number>(dst, src, F.P.data(), F.degree + 1, C.degree + 1,
                                        F.grid->n_cells(), s);
  else if(F.dg)
    e = launch_h_prolongate_dg_add<Number>(dst, src, F.P.data(), F.degree + 1, F.grid->topology(),
                                           C.grid->topology(), s);
  else
  {
    // separable passes x, y, z: coarse -> (f0,c1,c2) -> (f0,f1,c2) -> fine
    CGTopology const & qf = *F.op->cg_topology();
    CGTopology const & qc = *C.op->cg_topology();
    int const          d0[3] = {qc.nn[0], qc.nn[1], qc.nn[2]};
    int const          d1[3] = {qf.nn[0], qc.nn[1], qc.nn[2]};
    int const          d2[3] = {qf.nn[0], qf.nn[1], qc.nn[2]};
    Number *           wa = const_cast<Number *>(F.work_a.data());
    Number *           wb = const_cast<Number *>(F.work_b.data());
    e = launch_line_apply<Number>(wa, src, d0, 0, F.line_prolong[0], &qc, nullptr, false, s);
    if(e == hipSuccess)
      e = launch_line_apply<Number>(wb, wa, d1, 1, F.line_prolong[1], nullptr, nullptr, false, s);
    if(e == hipSuccess)
      e = launch_line_apply<Number>(dst, wb, d2, 2, F.line_prolong[2], nullptr, &qf, true, s);
  }
  if(e != hipSuccess)
    EXADG_THROW(EXADG_ERR_HIP, hipGetErrorString(e));
}

template<typename Number>
void MultigridPreconditioner<Number>::restrict_and_add(int l, Number * dst, Number const * src,
                                                       hipStream_t s) const
{
  Level const & F = levels_[l];
  Level const & C = levels_[l - 1];
  hipError_t    e = hipSuccess;
  if(F.transfer == 'c')
    e = launch_c_restrict_add<Number>(dst, src, F.grid->topology(), *C.op->cg_topology(), s);
  else if(F.dg && F.transfer == 'p')
    e = launch_p_restrict_add<Number>(dst, src, F.P.data(), F.degree + 1, C.degree + 1,
                                      F.grid->n_cells(), s);
  else if(F.dg)
    e = launch_h_restrict_dg_add<Number>(dst, src, F.P.data(), F.degree + 1, F.grid->topology(),
                                         C.grid->topology(), s);
  else
  {
    // transposed passes x, y, z: fine -> (c0,f1,f2) -> (c0,c1,f2) -> coarse
    CGTopology const & qf = *F.op->cg_topology();
    CGTopology const & qc = *C.op->cg_topology();
    int const          d0[3] = {qf.nn[0], qf.nn[1], qf.nn[2]};
    int const          d1[3] = {qc.nn[0], qf.nn[1], qf.nn[2]};
    int const          d2[3] = {qc.nn[0], qc.nn[1], qf.nn[2]};
    Number *           wa = const_cast<Number *>(F.work_a.data());
    Number *           wb = const_cast<Number *>(F.work_b.data());
    e = launch_line_apply<Number>(wa, src, d0, 0, F.line_restrict[0], &qf, nullptr, false, s);
    if(e == hipSuccess)
      e = launch_line_apply<Number>(wb, wa, d1, 1, F.line_restrict[1], nullptr, nullptr, false, s);
    if(e == hipSuccess)
      e = launch_line_apply<Number>(dst, wb, d2, 2, F.line_restrict[2], nullptr, &qc, true, s);
  }
  if(e != hipSuccess)
    EXADG_THROW(EXADG_ERR_HIP, hipGetErrorString(e));
}

template<typename Number>
void MultigridPreconditioner<Number>::coarse_solve(hipStream_t s)
{
  Level &       C = levels_[0];
  int64_t const n = C.op->n();
  if(data_.coarse_solver == CoarseSolver::chebyshev)
  {
    coarse_cheb_.vmult(C.sol, C.sol_alt, C.defect.data(), s);
    last_coarse_its_ = coarse_cheb_.degree;
    return;
  }
  // MGCoarseKrylov: r = src (mean removed if singular), PCG(Jacobi) from zero to
  // max(1e-12, rel_tol ||r0||)
  vec::copy(cr_.data(), C.defect.data(), n, s);
  if(data_.operator_is_singular)
    vec::subtract_mean(red_, cr_.data(), n, s);
  vec::dot(red_, 6, cr_.data(), cr_.data(), n, s);
  double const r0  = std::sqrt(red_.fetch(s, 7)[6]);
  double const tol = std::max(1e-12, data_.coarse_rel_tol * r0);
  JacobiData<Number> P;
  P.inv_diag       = coarse_inv_diag_.data();
  last_coarse_its_ = level_pcg<Number>(*C.op, P, cr_.data(), C.sol, tol,
                                       data_.coarse_max_iter, data_.operator_is_singular, red_, s,
                                       coarse_ws_);
}

template<typename Number>
void MultigridPreconditioner<Number>::v_cycle(int l, hipStream_t s)
{
  if(l == 0)
  {
    coarse_solve(s);
    return;
  }
  Level &       L = levels_[l];
  int64_t const n = L.op->n();
  L.smoother.vmult(L.sol, L.sol_alt, L.defect.data(), s);
  // t = defect - A sol, restricted to the next level's defect: one fused pass for a DG
  // p-transfer where the operator has it, else the residual kernel and the transfer
  bool const fused = L.dg && L.transfer == 'p' &&
                     L.op->residual_restrict(levels_[l - 1].defect.data(), L.sol, L.defect.data(),
                                             L.P.data(), levels_[l - 1].degree + 1, s);
  if(!fused)
  {
    L.op->residual(L.t.data(), L.sol, L.defect.data(), s);
    restrict_and_add(l, levels_[l - 1].defect.data(), L.t.data(), s);
  }
  v_cycle(l - 1, s);
  prolongate_and_add(l, L.sol, levels_[l - 1].sol, s);
  L.smoother.step(L.sol, L.sol_alt, L.defect.data(), s);
}

template<typename Number>
void MultigridPreconditioner<Number>::vmult_eager(double * dst, double const * src, hipStream_t s)
{
  int const     top = n_levels() - 1;
  int64_t const n   = levels_[top].op->n();
  if(src)
  {
    if constexpr(std::is_same<Number, double>::value)
      vec::copy(levels_[top].defect.data(), src, n, s);
    else
      vec::convert<Number, double>(levels_[top].defect.data(), src, n, s);
  }
  for(int l = 0; l < top; ++l)
    vec::set_zero(levels_[l].defect.data(), levels_[l].op->n(), s);
  v_cycle(top, s);
  if(dst)
  {
    if constexpr(std::is_same<Number, double>::value)
      vec::copy(dst, levels_[top].sol, n, s);
    else
      vec::convert<double, Number>(dst, levels_[top].sol, n, s);
  }
}

template<typename Number>
bool MultigridPreconditioner<Number>::smoother_vmult(double * dst, double const * src, hipStream_t s)
{
  Level &       L = levels_[n_levels() - 1];
  int64_t const n = L.op->n();
  if constexpr(std::is_same<Number, double>::value)
    vec::copy(L.defect.data(), src, n, s);
  else
    vec::convert<Number, double>(L.defect.data(), src, n, s);
  L.smoother.vmult(L.sol, L.sol_alt, L.defect.data(), s);
  if constexpr(std::is_same<Number, double>::value)
    vec::copy(dst, L.sol, n, s);
  else
    vec::convert<double, Number>(dst, L.sol, n, s);
  return true;
}

template<typename Number>
void MultigridPreconditioner<Number>::vmult(double * dst, double const * src, hipStream_t s)
{
  if(dst == nullptr || src == nullptr)
    EXADG_THROW(EXADG_ERR_ARGUMENT, "null vector");
  run(dst, src, s);
}

template<typename Number>
float * MultigridPreconditioner<Number>::narrow_input()
{
  if constexpr(std::is_same<Number, float>::value)
    return levels_.back().defect.data();
  else
    return nullptr;
}

template<typename Number>
float const * MultigridPreconditioner<Number>::vmult_narrow(hipStream_t s)
{
  if constexpr(std::is_same<Number, float>::value)
    return run(nullptr, nullptr, s);
  else
    return nullptr;
}

template<typename Number>
Number const * MultigridPreconditioner<Number>::run(double * dst, double const * src, hipStream_t s)
{
  // The first call runs eagerly (kernel attributes get set outside any capture). Later calls
  // replay a graph of the whole V-cycle: its ~10^2 small launches on the coarse levels cost
  // more in launch latency than in execution. Every buffer a cycle reads is written earlier in
  // the same cycle (or, with src = NULL, is the finest defect), so each replay computes the
  // same function of its input. The capture records only kernels and device copies (no memset
  // nodes) on a private stream; the replay runs on s. The smoothers swap the solution vectors
  // on the host, so each graph keeps the finest solution vector it ends in.
  if(!uses_graph() || eager_calls_ == 0)
  {
    vmult_eager(dst, src, s);
    ++eager_calls_;
    return levels_.back().sol;
  }
  Graph const * g = nullptr;
  for(Graph const & c : graphs_)
    if(c.dst == dst && c.src == src)
      g = &c;
  if(g == nullptr)
  {
    if(graphs_.size() >= 8)
      clear_graphs(), eager_calls_ = 1;
    if(capture_stream_ == nullptr)
      HIP_CHECK(hipStreamCreateWithFlags(&capture_stream_, hipStreamNonBlocking));
    hipGraph_t graph = nullptr;
    HIP_CHECK(hipStreamBeginCapture(capture_stream_, hipStreamCaptureModeThreadLocal));
    try
    {
      vmult_eager(dst, src, capture_stream_);
    }
    catch(...)
    {
      (void)hipStreamEndCapture(capture_stream_, &graph);
      if(graph)
        (void)hipGraphDestroy(graph);
      throw;
    }
    HIP_CHECK(hipStreamEndCapture(capture_stream_, &graph));
    hipGraphExec_t exec = nullptr;
    hipError_t const e  = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if(e != hipSuccess)
      EXADG_THROW(EXADG_ERR_HIP, hipGetErrorString(e));
    graphs_.push_back({dst, src, exec, levels_.back().sol});
    g = &graphs_.back();
  }
  HIP_CHECK(hipGraphLaunch(g->exec, s));
  return g->out;
}

std::unique_ptr<PreconditionerBase> create_multigrid(std::shared_ptr<StructuredGrid const> grid,
                                                     MultigridData const & data, double ip,
                                                     int precision, hipStream_t s)
{
  if(precision == EXADG_FP32)
    return std::make_unique<MultigridPreconditioner<float>>(std::move(grid), data, ip, s);
  if(precision == EXADG_FP64)
    return std::make_unique<MultigridPreconditioner<double>>(std::move(grid), data, ip, s);
  EXADG_THROW(EXADG_ERR_ARGUMENT, "multigrid precision must be EXADG_FP64 or EXADG_FP32");
}

template class LevelOperator<double>;
template class LevelOperator<float>;
template struct JacobiData<double>;
template struct JacobiData<float>;
template class Q1Operator<double>;
template class Q1Operator<float>;
template class CGOperator<double>;
template class CGOperator<float>;
template class ChebyshevSmoother<double>;
template class ChebyshevSmoother<float>;
template class MultigridPreconditioner<double>;
template class MultigridPreconditioner<float>;

}  // namespace ExaDG
