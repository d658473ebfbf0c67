#include "multigrid.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>

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

namespace
{
// host-synchronous PCG with point Jacobi on a level operator, recording the CG coefficients
// (setup-time eigenvalue estimates; coarse solves)
template<typename Number>
int level_pcg(LevelOperator<Number> const & op, Number const * inv_diag, Number const * b, Number * x,
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
  vec::scale(z.data(), inv_diag, r.data(), n, s);
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
    vec::scale(z.data(), inv_diag, r.data(), n, s);
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
  void residual(Number * dst, Number const * src, Number const * b, hipStream_t s) const override
  {
    op_->residual(dst, src, b, s);
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

// ---------------------------------------------------------------- Chebyshev

template<typename Number>
void ChebyshevSmoother<Number>::setup(LevelOperator<Number> const & op, int degree_,
                                      double smoothing_range, int eig_cg_iterations, hipStream_t s)
{
  op_          = &op;
  degree       = degree_;
  int64_t const n = op.n();
  inv_diag_.reinit(n);
  t_.reinit(n);
  op.inverse_diagonal(inv_diag_.data(), s);
  // start vector of deal.II's eigenvalue estimate: (i mod 11) minus its mean, constrained
  // entries zero; eig_cg_iterations steps of Jacobi-PCG, Lanczos tridiagonal
  std::vector<Number> h((size_t)n);
  double              mean = 0;
  for(int64_t i = 0; i < n; ++i)
    mean += (double)(i % 11);
  mean /= (double)n;
  std::vector<char> const mask = op.constrained_mask();
  double                  nrm2 = 0;
  for(int64_t i = 0; i < n; ++i)
  {
    double v = (double)(i % 11) - mean;
    if(!mask.empty() && mask[i])
      v = 0;
    h[i] = (Number)v;
    nrm2 += v * v;
  }
  DeviceVector<Number> b(n), x(n);
  HIP_CHECK(hipMemcpyAsync(b.data(), h.data(), n * sizeof(Number), hipMemcpyHostToDevice, s));
  Reducer              red;
  PcgWorkspace<Number> ws;
  std::vector<double>  al, be;
  cg_iterations = level_pcg<Number>(op, inv_diag_.data(), b.data(), x.data(), 1e-2 * std::sqrt(nrm2),
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
  cg_iterations = level_pcg<Number>(op, inv_diag_.data(), b.data(), x.data(), 1e-5 * std::sqrt(nrm2),
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
  if(op_->chebyshev_step(x_new, x, x_old, b, inv_diag_.data(), f1, f2, has_old, s))
    return;
  op_->vmult(t_.data(), x, s);
  vec::chebyshev_update<Number>(x_new, x, x_old, b, t_.data(), inv_diag_.data(), f1, f2, false,
                                has_old, op_->n(), s);
}

// ---------------------------------------------------------------- multigrid

template<typename Number>
MultigridPreconditioner<Number>::MultigridPreconditioner(std::shared_ptr<StructuredGrid const> grid,
                                                         MultigridData const & data, double ip,
                                                         hipStream_t s)
  : grid_(std::move(grid)), data_(data)
{
  if(data.degrees.empty())
    EXADG_THROW(EXADG_ERR_ARGUMENT, "multigrid needs at least one DG level");
  for(size_t i = 1; i < data.degrees.size(); ++i)
    if(data.degrees[i] >= data.degrees[i - 1] || data.degrees[i] < 1)
      EXADG_THROW(EXADG_ERR_ARGUMENT, "DG degrees must decrease strictly from fine to coarse");
  if(data.coarse_q1 && data.degrees.back() != 1)
    EXADG_THROW(EXADG_ERR_ARGUMENT, "the continuous Q1 level must follow a DG degree-1 level");
  // levels coarse (0) to fine
  int const n_levels = (int)data.degrees.size() + (data.coarse_q1 ? 1 : 0);
  levels_.resize(n_levels);
  int l = 0;
  if(data.coarse_q1)
  {
    levels_[0].op     = std::make_unique<Q1Operator<Number>>(*grid_);
    levels_[0].degree = 1;
    levels_[0].dg     = false;
    ++l;
  }
  for(int i = (int)data.degrees.size() - 1; i >= 0; --i, ++l)
  {
    levels_[l].op     = std::make_unique<DGLevelOperator<Number>>(grid_, data.degrees[i], ip);
    levels_[l].degree = data.degrees[i];
    levels_[l].dg     = true;
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
    if(k > 0 && L.dg && levels_[k - 1].dg)
    {
      // p-transfer: P[a][i] = phi^coarse_i(x^fine_a) on Gauss-Lobatto nodes
      int const           nf = L.degree + 1, nc = levels_[k - 1].degree + 1;
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
                              data_.eig_cg_iterations, s);
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
  hipError_t    e;
  if(C.dg)
    e = launch_p_prolongate_add<Number>(dst, src, F.P.data(), F.degree + 1, C.degree + 1,
                                        grid_->n_cells(), s);
  else
    e = launch_c_prolongate_add<Number>(dst, src, grid_->topology(),
                                        static_cast<Q1Operator<Number> const &>(*C.op).q1_topology(), s);
  if(e != hipSuccess)
    EXADG_THROW(EXADG_ERR_HIP, hipGetErrorString(e));
}

template<typename Number>
void MultigridPreconditioner<Number>::restrict_and_add(int l, Number * dst, Number const * src,
                                                       hipStream_t s) const
{
  Level const & F = levels_[l];
  Level const & C = levels_[l - 1];
  hipError_t    e;
  if(C.dg)
    e = launch_p_restrict_add<Number>(dst, src, F.P.data(), F.degree + 1, C.degree + 1,
                                      grid_->n_cells(), s);
  else
    e = launch_c_restrict_add<Number>(dst, src, grid_->topology(),
                                      static_cast<Q1Operator<Number> const &>(*C.op).q1_topology(), s);
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
  last_coarse_its_ = level_pcg<Number>(*C.op, coarse_inv_diag_.data(), cr_.data(), C.sol, tol,
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
  L.op->residual(L.t.data(), L.sol, L.defect.data(), s);  // t = defect - A sol
  restrict_and_add(l, levels_[l - 1].defect.data(), L.t.data(), s);
  v_cycle(l - 1, s);
  prolongate_and_add(l, L.sol, levels_[l - 1].sol, s);
  L.smoother.step(L.sol, L.sol_alt, L.defect.data(), s);
}

template<typename Number>
void MultigridPreconditioner<Number>::vmult_eager(double * dst, double const * src, hipStream_t s)
{
  int const     top = n_levels() - 1;
  int64_t const n   = levels_[top].op->n();
  if constexpr(std::is_same<Number, double>::value)
    vec::copy(levels_[top].defect.data(), src, n, s);
  else
    vec::convert<Number, double>(levels_[top].defect.data(), src, n, s);
  for(int l = 0; l < top; ++l)
    vec::set_zero(levels_[l].defect.data(), levels_[l].op->n(), s);
  v_cycle(top, s);
  if constexpr(std::is_same<Number, double>::value)
    vec::copy(dst, levels_[top].sol, n, s);
  else
    vec::convert<double, Number>(dst, levels_[top].sol, n, s);
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
  // The first call runs eagerly (kernel attributes get set outside any capture). Later calls
  // replay a graph of the whole V-cycle: its ~10^2 small launches on the coarse levels cost
  // more in launch latency than in execution. Every buffer a cycle reads is written earlier in
  // the same cycle, so each replay computes the same function of src. The capture records
  // only kernels and device copies (no memset nodes) on a private stream; the replay runs on s.
  if(!uses_graph() || eager_calls_ == 0)
  {
    vmult_eager(dst, src, s);
    ++eager_calls_;
    return;
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
    graphs_.push_back({dst, src, exec});
    g = &graphs_.back();
  }
  HIP_CHECK(hipGraphLaunch(g->exec, s));
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
template class Q1Operator<double>;
template class Q1Operator<float>;
template class ChebyshevSmoother<double>;
template class ChebyshevSmoother<float>;
template class MultigridPreconditioner<double>;
template class MultigridPreconditioner<float>;

}  // namespace ExaDG
