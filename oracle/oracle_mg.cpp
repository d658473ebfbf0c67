// CPU oracle multigrid (TEST INFRASTRUCTURE ONLY). See oracle_mg.h for the reference files.
#include "oracle_mg.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <stdexcept>

namespace oracle
{
// ---------------------------------------------------------------- eigenvalues

void tridiagonal_extreme_eigenvalues(std::vector<double> const & alpha,
                                     std::vector<double> const & beta, double & lmin, double & lmax)
{
  // Lanczos matrix of CG, eigenvalues by the implicit QL algorithm with Wilkinson shifts
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
  for(int l = 0; l < n; ++l)
  {
    int iter = 0, m;
    do
    {
      for(m = l; m < n - 1; ++m)
      {
        double const dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if(std::fabs(e[m]) <= 1e-15 * dd)
          break;
      }
      if(m != l)
      {
        if(++iter > 200)
          throw std::runtime_error("QL did not converge");
        double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
        double r = std::hypot(g, 1.0);
        g        = d[m] - d[l] + e[l] / (g + (g >= 0 ? std::fabs(r) : -std::fabs(r)));
        double s = 1.0, c = 1.0, p = 0.0;
        int    i;
        for(i = m - 1; i >= l; --i)
        {
          double f = s * e[i], b = c * e[i];
          e[i + 1] = (r = std::hypot(f, g));
          if(r == 0.0)
          {
            d[i + 1] -= p;
            e[m] = 0.0;
            break;
          }
          s = f / r;
          c = g / r;
          g = d[i + 1] - p;
          r = (d[i] - g) * s + 2.0 * c * b;
          d[i + 1] = g + (p = s * r);
          g        = c * r - b;
        }
        if(r == 0.0 && i >= l)
          continue;
        d[l] -= p;
        e[l] = g;
        e[m] = 0.0;
      }
    } while(m != l);
  }
  lmin = *std::min_element(d.begin(), d.end());
  lmax = *std::max_element(d.begin(), d.end());
}

static double dotv(int64_t n, double const * a, double const * b)
{
  double s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for(int64_t i = 0; i < n; ++i)
    s += a[i] * b[i];
  return s;
}

static void subtract_mean(std::vector<double> & v)
{
  double s = 0;
  for(double x : v)
    s += x;
  s /= (double)v.size();
  for(double & x : v)
    x -= s;
}

int pcg_record(int64_t n, Apply const & A, Apply const & P, double const * b, double * x,
               double tol_abs, int max_iter, bool singular, std::vector<double> * alphas,
               std::vector<double> * betas)
{
  // zero initial guess
  std::vector<double> r(b, b + n), z(n), p(n), v(n);
  std::fill(x, x + n, 0.0);
  if(singular)
    subtract_mean(r);
  P(z.data(), r.data());
  double rho = dotv(n, r.data(), z.data());
  if(std::sqrt(dotv(n, r.data(), r.data())) <= tol_abs)
    return 0;
  p      = z;
  int it = 0;
  for(it = 1; it <= max_iter; ++it)
  {
    A(v.data(), p.data());
    double const alpha = rho / dotv(n, p.data(), v.data());
    for(int64_t i = 0; i < n; ++i)
    {
      x[i] += alpha * p[i];
      r[i] -= alpha * v[i];
    }
    if(singular)
      subtract_mean(r);
    P(z.data(), r.data());
    double const rho_new = dotv(n, r.data(), z.data());
    double const rn      = std::sqrt(dotv(n, r.data(), r.data()));
    double const beta    = rho_new / rho;
    if(alphas)
      alphas->push_back(alpha);
    if(betas)
      betas->push_back(beta);
    if(rn <= tol_abs)
      break;
    rho = rho_new;
    for(int64_t i = 0; i < n; ++i)
      p[i] = z[i] + beta * p[i];
  }
  return std::min(it, max_iter);
}

// ---------------------------------------------------------------- Q1 operator

Q1Operator::Q1Operator(MeshDesc const & d) : mesh(d)
{
  for(int k = 0; k < 3; ++k)
    nn[k] = d.n_cells[k] + (d.periodic[k] ? 0 : 1);
  n_nodes = (int64_t)nn[0] * nn[1] * nn[2];
  constrained.assign(n_nodes, 0);
  for(int64_t i = 0; i < n_nodes; ++i)
  {
    int const c[3] = {(int)(i % nn[0]), (int)((i / nn[0]) % nn[1]), (int)(i / ((int64_t)nn[0] * nn[1]))};
    for(int k = 0; k < 3; ++k)
      if(!d.periodic[k] && ((c[k] == 0 && d.boundary_type[2 * k] == 0) ||
                            (c[k] == d.n_cells[k] && d.boundary_type[2 * k + 1] == 0)))
        constrained[i] = 1;
  }
  std::vector<double> gp, gw;
  gauss_legendre(2, gp, gw);
  K.assign((size_t)mesh.n_cells * 64, 0.0);
#pragma omp parallel for schedule(static)
  for(int64_t c = 0; c < mesh.n_cells; ++c)
  {
    double * Kc = &K[(size_t)c * 64];
    for(int q = 0; q < 8; ++q)
    {
      double const xi[3] = {gp[q & 1], gp[(q >> 1) & 1], gp[(q >> 2) & 1]};
      double const w     = gw[q & 1] * gw[(q >> 1) & 1] * gw[(q >> 2) & 1];
      double       x[3], J[9], Ji[9], det;
      mesh.map_point(c, xi, x, J);
      invert3(J, Ji, det);
      double g[8][3];
      for(int a = 0; a < 8; ++a)
      {
        double gr[3];
        for(int e = 0; e < 3; ++e)
        {
          double v = 1;
          for(int k = 0; k < 3; ++k)
          {
            int const bit = (a >> k) & 1;
            v *= (k == e) ? (bit ? 1.0 : -1.0) : (bit ? xi[k] : 1.0 - xi[k]);
          }
          gr[e] = v;
        }
        for(int i = 0; i < 3; ++i)
          g[a][i] = Ji[0 * 3 + i] * gr[0] + Ji[1 * 3 + i] * gr[1] + Ji[2 * 3 + i] * gr[2];
      }
      for(int a = 0; a < 8; ++a)
        for(int b = 0; b < 8; ++b)
          Kc[a * 8 + b] += det * w * (g[a][0] * g[b][0] + g[a][1] * g[b][1] + g[a][2] * g[b][2]);
    }
  }
}

int64_t Q1Operator::node_of(int const * cc, int v) const
{
  int nd[3];
  for(int k = 0; k < 3; ++k)
  {
    int c = cc[k] + ((v >> k) & 1);
    if(mesh.d.periodic[k] && c == mesh.d.n_cells[k])
      c = 0;
    nd[k] = c;
  }
  return nd[0] + (int64_t)nn[0] * (nd[1] + (int64_t)nn[1] * nd[2]);
}

void Q1Operator::vmult(double * dst, double const * src) const
{
  // cell-wise assembly of the action (serial scatter), constrained rows identity
  std::fill(dst, dst + n_nodes, 0.0);
  for(int64_t c = 0; c < mesh.n_cells; ++c)
  {
    int cc[3];
    mesh.cell_coords(c, cc[0], cc[1], cc[2]);
    int64_t nodes[8];
    double  u[8];
    for(int v = 0; v < 8; ++v)
    {
      nodes[v] = node_of(cc, v);
      u[v]     = constrained[nodes[v]] ? 0.0 : src[nodes[v]];
    }
    double const * Kc = &K[(size_t)c * 64];
    for(int a = 0; a < 8; ++a)
    {
      double s = 0;
      for(int b = 0; b < 8; ++b)
        s += Kc[a * 8 + b] * u[b];
      dst[nodes[a]] += s;
    }
  }
  for(int64_t i = 0; i < n_nodes; ++i)
    if(constrained[i])
      dst[i] = src[i];
}

void Q1Operator::diagonal(double * dg) const
{
  std::fill(dg, dg + n_nodes, 0.0);
  for(int64_t c = 0; c < mesh.n_cells; ++c)
  {
    int cc[3];
    mesh.cell_coords(c, cc[0], cc[1], cc[2]);
    for(int a = 0; a < 8; ++a)
      dg[node_of(cc, a)] += K[(size_t)c * 64 + a * 9];
  }
  for(int64_t i = 0; i < n_nodes; ++i)
    if(constrained[i])
      dg[i] = 1.0;
}

// ---------------------------------------------------------------- continuous Q_p levels

CGSpace::CGSpace(MeshDesc const & md, int p_) : d(md), p(p_)
{
  for(int k = 0; k < 3; ++k)
    nn[k] = d.n_cells[k] * p + (d.periodic[k] ? 0 : 1);
  n_nodes = (int64_t)nn[0] * nn[1] * nn[2];
  constrained.assign(n_nodes, 0);
  for(int64_t g = 0; g < n_nodes; ++g)
  {
    int const c[3] = {(int)(g % nn[0]), (int)((g / nn[0]) % nn[1]), (int)(g / ((int64_t)nn[0] * nn[1]))};
    for(int k = 0; k < 3; ++k)
      if(!d.periodic[k] && ((c[k] == 0 && d.boundary_type[2 * k] == 0) ||
                            (c[k] == nn[k] - 1 && d.boundary_type[2 * k + 1] == 0)))
        constrained[g] = 1;
  }
}

int64_t CGSpace::node(int64_t c, int local) const
{
  int const n = p + 1;
  int64_t   cc[3];
  cc[0] = c % d.n_cells[0];
  cc[1] = (c / d.n_cells[0]) % d.n_cells[1];
  cc[2] = c / ((int64_t)d.n_cells[0] * d.n_cells[1]);
  int const loc[3] = {local % n, (local / n) % n, local / (n * n)};
  int64_t   g[3];
  for(int k = 0; k < 3; ++k)
  {
    g[k] = cc[k] * p + loc[k];
    if(d.periodic[k] && g[k] == nn[k])
      g[k] = 0;
  }
  return g[0] + (int64_t)nn[0] * (g[1] + (int64_t)nn[1] * g[2]);
}

int CGSpace::valence(int64_t g) const
{
  int64_t const c[3] = {g % nn[0], (g / nn[0]) % nn[1], g / ((int64_t)nn[0] * nn[1])};
  int           v    = 1;
  for(int k = 0; k < 3; ++k)
  {
    bool const vertex = c[k] % p == 0;
    bool const end    = !d.periodic[k] && (c[k] == 0 || c[k] == nn[k] - 1);
    if(vertex && !end)
      v *= 2;
  }
  return v;
}

CGOperator::CGOperator(MeshDesc const & d, int p, double ip)
  : space(d, p), cells(std::make_unique<LaplaceOperator>(d, p, ip))
{
}

void CGOperator::vmult(double * dst, double const * src) const
{
  // OperatorBase on a continuous DoFHandler: gather (constrained entries zero), cell integrals,
  // assembly; constrained rows act as the identity (operator_base.cpp:299-309)
  int64_t const       nc = space.d.n_cells[0] * (int64_t)space.d.n_cells[1] * space.d.n_cells[2];
  int const           n3 = (space.p + 1) * (space.p + 1) * (space.p + 1);
  std::vector<double> u((size_t)nc * n3), v((size_t)nc * n3, 0.0);
  for(int64_t c = 0; c < nc; ++c)
    for(int i = 0; i < n3; ++i)
    {
      int64_t const g = space.node(c, i);
      u[c * n3 + i]   = space.constrained[g] ? 0.0 : src[g];
    }
  cells->cell_loop(v.data(), u.data());
  std::fill(dst, dst + space.n_nodes, 0.0);
  for(int64_t c = 0; c < nc; ++c)
    for(int i = 0; i < n3; ++i)
      dst[space.node(c, i)] += v[c * n3 + i];
  for(int64_t g = 0; g < space.n_nodes; ++g)
    if(space.constrained[g])
      dst[g] = src[g];
}

void CGOperator::diagonal(double * dg) const
{
  // local diagonal by probing the cell integral with unit vectors (one per local index,
  // in all cells at once), summed over the cells sharing a node
  int64_t const       nc = space.d.n_cells[0] * (int64_t)space.d.n_cells[1] * space.d.n_cells[2];
  int const           n3 = (space.p + 1) * (space.p + 1) * (space.p + 1);
  std::vector<double> e((size_t)nc * n3), out((size_t)nc * n3);
  std::fill(dg, dg + space.n_nodes, 0.0);
  for(int i = 0; i < n3; ++i)
  {
    std::fill(e.begin(), e.end(), 0.0);
    std::fill(out.begin(), out.end(), 0.0);
    for(int64_t c = 0; c < nc; ++c)
      e[c * n3 + i] = 1.0;
    cells->cell_loop(out.data(), e.data());
    for(int64_t c = 0; c < nc; ++c)
      dg[space.node(c, i)] += out[c * n3 + i];
  }
  for(int64_t g = 0; g < space.n_nodes; ++g)
    if(space.constrained[g])
      dg[g] = 1.0;
}

// ---------------------------------------------------------------- block Jacobi

BlockJacobi::BlockJacobi(LaplaceOperator const & op) : m(op.n3), n_cells(op.mesh.n_cells)
{
  L.resize((size_t)n_cells * m * m);
  op.block_diagonal(L.data());
#pragma omp parallel for schedule(static)
  for(int64_t c = 0; c < n_cells; ++c)
  {
    double * A = &L[(size_t)c * m * m];  // symmetric: column storage = row storage
    for(int j = 0; j < m; ++j)
    {
      double d = A[j * m + j];
      for(int k = 0; k < j; ++k)
        d -= A[j * m + k] * A[j * m + k];
      if(!(d > 0))
        throw std::runtime_error("block Jacobi: cell block not positive definite");
      double const ljj = std::sqrt(d);
      A[j * m + j]     = ljj;
      for(int i = j + 1; i < m; ++i)
      {
        double v = A[i * m + j];
        for(int k = 0; k < j; ++k)
          v -= A[i * m + k] * A[j * m + k];
        A[i * m + j] = v / ljj;
      }
    }
  }
}

void BlockJacobi::vmult(double * dst, double const * src) const
{
#pragma omp parallel for schedule(static)
  for(int64_t c = 0; c < n_cells; ++c)
  {
    double const * A = &L[(size_t)c * m * m];
    double *       y = dst + c * m;
    double const * r = src + c * m;
    for(int i = 0; i < m; ++i)  // L y = r
    {
      double v = r[i];
      for(int k = 0; k < i; ++k)
        v -= A[i * m + k] * y[k];
      y[i] = v / A[i * m + i];
    }
    for(int i = m - 1; i >= 0; --i)  // L^T x = y
    {
      double v = y[i];
      for(int k = i + 1; k < m; ++k)
        v -= A[k * m + i] * y[k];
      y[i] = v / A[i * m + i];
    }
  }
}

// ---------------------------------------------------------------- Chebyshev

static Apply jacobi(std::vector<double> const & inv_diag)
{
  return [&inv_diag](double * dst, double const * src) {
    int64_t const n = (int64_t)inv_diag.size();
    for(int64_t i = 0; i < n; ++i)
      dst[i] = inv_diag[i] * src[i];
  };
}

void Chebyshev::setup_smoother(Apply const & A_, std::vector<double> dinv,
                               std::vector<char> const & mask, int degree_, double range,
                               int eig_its, Apply const & P_)
{
  A        = A_;
  inv_diag = std::move(dinv);
  P        = P_;
  degree   = degree_;
  int64_t const n = (int64_t)inv_diag.size();
  // deal.II initial vector: (i mod 11) minus mean (entries in `mask` zero; the multigrid passes
  // none, as ExaDG's smoother sets no constraints object)
  std::vector<double> b(n), x(n);
  for(int64_t i = 0; i < n; ++i)
    b[i] = (double)(i % 11);
  subtract_mean(b);
  for(int64_t i = 0; i < n; ++i)
    if(!mask.empty() && mask[i])
      b[i] = 0;
  std::vector<double> al, be;
  double const        nb = std::sqrt(dotv(n, b.data(), b.data()));
  cg_iterations = pcg_record(n, A, P ? P : jacobi(inv_diag), b.data(), x.data(), 1e-2 * nb,
                             eig_its, false, &al, &be);
  tridiagonal_extreme_eigenvalues(al, be, lambda_min, lambda_max);
  double const max_ev = 1.2 * lambda_max;
  double const lo     = max_ev / range;
  theta               = 0.5 * (max_ev + lo);
  delta               = 0.5 * (max_ev - lo);
}

void Chebyshev::setup_coarse(Apply const & A_, std::vector<double> dinv, double rel_tol,
                             bool singular)
{
  A                  = A_;
  inv_diag           = std::move(dinv);
  int64_t const n    = (int64_t)inv_diag.size();
  std::vector<double> b(n), x(n);
  srand(1);
  for(int64_t i = 0; i < n; ++i)
    b[i] = (double)rand() / RAND_MAX;
  if(singular)
    subtract_mean(b);
  std::vector<double> al, be;
  double const        nb = std::sqrt(dotv(n, b.data(), b.data()));
  cg_iterations = pcg_record(n, A, jacobi(inv_diag), b.data(), x.data(), 1e-5 * nb, 10000, singular,
                             &al, &be);
  tridiagonal_extreme_eigenvalues(al, be, lambda_min, lambda_max);
  double const factor = 1.1;
  double const max_ev = factor * lambda_max;
  double const range  = lambda_max / lambda_min * factor;
  double const sigma  = (1. - std::sqrt(1. / range)) / (1. + std::sqrt(1. / range));
  degree = std::max(1, (int)(std::log(1. / rel_tol + std::sqrt(1. / rel_tol / rel_tol - 1.)) /
                             std::log(1. / sigma)));
  double const lo = max_ev / range;
  theta           = 0.5 * (max_ev + lo);
  delta           = 0.5 * (max_ev - lo);
}

// z = P^{-1} r with the smoother's inner preconditioner
static void precondition(Chebyshev const & c, std::vector<double> & z, std::vector<double> const & r)
{
  if(c.P)
    c.P(z.data(), r.data());
  else
    for(size_t i = 0; i < r.size(); ++i)
      z[i] = c.inv_diag[i] * r[i];
}

void Chebyshev::vmult(std::vector<double> & x, std::vector<double> const & b) const
{
  // deal.II PreconditionChebyshev::vmult: zero initial guess
  int64_t const       n = (int64_t)b.size();
  std::vector<double> xo(n, 0.0), t(n), r(n), z(n);
  precondition(*this, z, b);
  for(int64_t i = 0; i < n; ++i)
    x[i] = z[i] / theta;
  if(degree < 2 || std::fabs(delta) < 1e-40)
    return;
  double rhok = delta / theta, sigma = theta / delta;
  for(int it = 1; it < degree; ++it)
  {
    A(t.data(), x.data());
    for(int64_t i = 0; i < n; ++i)
      r[i] = b[i] - t[i];
    precondition(*this, z, r);
    double const rhokp = 1. / (2. * sigma - rhok);
    double const f1 = rhokp * rhok, f2 = 2. * rhokp / delta;
    rhok            = rhokp;
    for(int64_t i = 0; i < n; ++i)
    {
      double const xn = (1. + f1) * x[i] - (it > 1 ? f1 * xo[i] : 0.0) + f2 * z[i];
      xo[i]           = x[i];
      x[i]            = xn;
    }
  }
}

void Chebyshev::step(std::vector<double> & x, std::vector<double> const & b) const
{
  int64_t const       n = (int64_t)b.size();
  std::vector<double> xo(n), t(n), r(n), z(n);
  A(t.data(), x.data());
  for(int64_t i = 0; i < n; ++i)
    r[i] = b[i] - t[i];
  precondition(*this, z, r);
  for(int64_t i = 0; i < n; ++i)
  {
    xo[i] = x[i];
    x[i]  = x[i] + z[i] / theta;
  }
  if(degree < 2 || std::fabs(delta) < 1e-40)
    return;
  double rhok = delta / theta, sigma = theta / delta;
  for(int it = 1; it < degree; ++it)
  {
    A(t.data(), x.data());
    for(int64_t i = 0; i < n; ++i)
      r[i] = b[i] - t[i];
    precondition(*this, z, r);
    double const rhokp = 1. / (2. * sigma - rhok);
    double const f1 = rhokp * rhok, f2 = 2. * rhokp / delta;
    rhok            = rhokp;
    for(int64_t i = 0; i < n; ++i)
    {
      double const xn = (1. + f1) * x[i] - f1 * xo[i] + f2 * z[i];
      xo[i]           = x[i];
      x[i]            = xn;
    }
  }
}

// ---------------------------------------------------------------- multigrid

std::vector<LevelInfo> mg_level_list(int type, std::vector<int> const & degrees, int n_h, bool coarse_q1)
{
  if(type == mg_legacy)
    type = coarse_q1 ? mg_pcMG : mg_pMG;
  auto has = [&](std::initializer_list<int> l) {
    for(int t : l)
      if(t == type)
        return true;
    return false;
  };
  bool const h_involved = has({mg_hMG, mg_chMG, mg_hcMG, mg_hpMG, mg_chpMG, mg_hcpMG, mg_hpcMG,
                               mg_phMG, mg_cphMG, mg_pchMG, mg_phcMG});
  int const  H          = h_involved ? std::max(1, n_h) : 1;
  int const  k          = degrees.at(0);
  // p-levels (degree, dg), built fine to coarse and then reversed
  std::vector<LevelInfo> pl;
  if(type == mg_hMG)
    pl.push_back({0, k, true});
  else if(has({mg_cMG, mg_chMG, mg_hcMG}))
  {
    pl.push_back({0, k, false});
    pl.push_back({0, k, true});
  }
  else
  {
    bool dg = true;
    if(has({mg_cpMG, mg_hcpMG, mg_chpMG, mg_cphMG}))  // c-transfer before p-coarsening
    {
      pl.push_back({0, k, true});
      dg = false;
    }
    for(int p : degrees)
      pl.push_back({0, p, dg});
    if(has({mg_pcMG, mg_hpcMG, mg_phcMG, mg_pchMG}))  // c-transfer after p-coarsening
      pl.push_back({0, degrees.back(), false});
    std::reverse(pl.begin(), pl.end());
  }
  std::vector<LevelInfo> L;
  auto at = [](int h, LevelInfo p) { return LevelInfo{h, p.degree, p.dg}; };
  int const P = (int)pl.size();
  if(type == mg_hMG)
    for(int h = 0; h < H; ++h)
      L.push_back(at(h, pl.front()));
  else if(type == mg_cMG)
  {
    L.push_back(at(H - 1, pl.front()));
    L.push_back(at(H - 1, pl.back()));
  }
  else if(type == mg_chMG)
  {
    for(int h = 0; h < H; ++h)
      L.push_back(at(h, pl.front()));
    L.push_back(at(H - 1, pl.back()));
  }
  else if(type == mg_hcMG)
  {
    L.push_back(at(0, pl.front()));
    for(int h = 0; h < H; ++h)
      L.push_back(at(h, pl.back()));
  }
  else if(has({mg_pMG, mg_pcMG, mg_cpMG}))
    for(auto const & p : pl)
      L.push_back(at(0, p));
  else if(has({mg_phMG, mg_cphMG, mg_pchMG}))
  {
    for(int h = 0; h + 1 < H; ++h)
      L.push_back(at(h, pl.front()));
    for(auto const & p : pl)
      L.push_back(at(H - 1, p));
  }
  else if(has({mg_hpMG, mg_hcpMG, mg_hpcMG}))
  {
    for(int i = 0; i + 1 < P; ++i)
      L.push_back(at(0, pl[i]));
    for(int h = 0; h < H; ++h)
      L.push_back(at(h, pl.back()));
  }
  else if(type == mg_phcMG)
  {
    L.push_back(at(0, pl.front()));
    for(int h = 0; h + 1 < H; ++h)
      L.push_back(at(h, pl[1]));
    for(int i = 1; i < P; ++i)
      L.push_back(at(H - 1, pl[i]));
  }
  else if(type == mg_chpMG)
  {
    for(int i = 0; i + 2 < P; ++i)
      L.push_back(at(0, pl[i]));
    for(int h = 0; h < H; ++h)
      L.push_back(at(h, pl[P - 2]));
    L.push_back(at(H - 1, pl.back()));
  }
  else
    throw std::runtime_error("unknown multigrid type");
  for(size_t l = 1; l < L.size(); ++l)
  {
    int const changes = (L[l].h != L[l - 1].h) + (L[l].degree != L[l - 1].degree) +
                        (L[l].dg != L[l - 1].dg);
    if(changes != 1)
      throw std::runtime_error("between two multigrid levels exactly one of h, p, c must change");
  }
  return L;
}

static MeshDesc coarsened(MeshDesc m, int times)
{
  for(int k = 0; k < 3; ++k)
  {
    if(m.n_cells[k] % (1 << times) != 0)
      throw std::runtime_error("h-levels need cell counts divisible by 2 per coarsening");
    m.n_cells[k] >>= times;
  }
  return m;
}

Multigrid::Multigrid(MeshDesc const & m, MGDesc const & d, double ip) : mesh(m), desc(d)
{
  std::vector<LevelInfo> const info = mg_level_list(d.type, d.degrees, d.n_h_levels, d.coarse_q1);
  int const                    nl   = (int)info.size();
  int const                    H    = info.back().h + 1;
  levels.resize(nl);
  for(int l = 0; l < nl; ++l)
  {
    Level & L = levels[l];
    L.h       = info[l].h;
    L.degree  = info[l].degree;
    L.dg      = info[l].dg;
    L.mesh    = coarsened(m, H - 1 - L.h);
    if(L.dg)
    {
      L.dg_op              = std::make_unique<LaplaceOperator>(L.mesh, L.degree, ip);
      L.n                  = L.dg_op->n_dofs();
      LaplaceOperator * op = L.dg_op.get();
      L.A                  = [op](double * dst, double const * src) { op->vmult(dst, src); };
      L.inv_diag.resize(L.n);
      op->diagonal(L.inv_diag.data());
    }
    else
    {
      L.space = CGSpace(L.mesh, L.degree);
      L.n     = L.space.n_nodes;
      L.inv_diag.resize(L.n);
      if(L.degree == 1)
      {
        L.q1_op        = std::make_unique<Q1Operator>(L.mesh);
        Q1Operator * q = L.q1_op.get();
        L.A            = [q](double * dst, double const * src) { q->vmult(dst, src); };
        q->diagonal(L.inv_diag.data());
      }
      else
      {
        L.cg_op        = std::make_unique<CGOperator>(L.mesh, L.degree, ip);
        CGOperator * q = L.cg_op.get();
        L.A            = [q](double * dst, double const * src) { q->vmult(dst, src); };
        q->diagonal(L.inv_diag.data());
      }
    }
    invert_diagonal(L.inv_diag);
  }
  for(int k = 0; k < nl; ++k)
  {
    Level & L = levels[k];
    L.defect.assign(L.n, 0.0);
    L.sol.assign(L.n, 0.0);
    L.t.assign(L.n, 0.0);
    if(k == 0)
      continue;
    std::vector<char> mask;
    Apply             P;
    if(d.smoother_preconditioner != 0)
    {
      if(!L.dg)
        throw std::runtime_error("block Jacobi / additive Schwarz smoothing needs DG levels");
      L.block           = std::make_unique<BlockJacobi>(*L.dg_op);
      BlockJacobi * blk = L.block.get();
      P                 = [blk](double * dst, double const * src) { blk->vmult(dst, src); };
    }
    L.smoother.setup_smoother(L.A, L.inv_diag, mask, d.smoother_degree, d.smoothing_range,
                              d.eig_its, P);
    Level const & C = levels[k - 1];
    if(C.dg != L.dg)
      L.transfer = 'c';
    else if(C.degree != L.degree)
    {
      L.transfer = 'p';
      int const           nf = L.degree + 1, nc = C.degree + 1;
      std::vector<double> xf, xc;
      gauss_lobatto(nf, xf);
      gauss_lobatto(nc, xc);
      L.P.resize(nf * nc);
      for(int a = 0; a < nf; ++a)
        for(int i = 0; i < nc; ++i)
          L.P[a * nc + i] = lagrange(xc, i, xf[a]);
    }
    else
    {
      L.transfer = 'h';
      int const           n = L.degree + 1;
      std::vector<double> x;
      gauss_lobatto(n, x);
      for(int b = 0; b < 2; ++b)
      {
        L.Ph[b].resize(n * n);
        for(int a = 0; a < n; ++a)
          for(int i = 0; i < n; ++i)
            L.Ph[b][a * n + i] = lagrange(x, i, 0.5 * (b + x[a]));
      }
    }
  }
  if(d.coarse_solver == 1)
    coarse_cheb.setup_coarse(levels[0].A, levels[0].inv_diag, d.coarse_rel_tol, d.singular);
}

// Cell-wise transfer between level l-1 (coarse) and l (fine). For every fine cell, the coarse
// values of its parent cell (the cell itself for 'p') are interpolated with the 1D matrices of
// the cell's position in the parent. On continuous levels the values are gathered from and
// scattered to global nodes: prolongation writes each fine node once (every cell sharing it
// computes the same value), restriction weights each fine node by 1/valence so that the sum over
// the cells sharing it counts it once -- restriction is then the transpose of prolongation.
// Constrained nodes read and receive zero.
namespace
{
struct CellTransfer
{
  int                 nf, nc;  // 1D sizes
  double const *      M[3];    // per direction nf x nc
};

void parent_and_matrices(Multigrid::Level const & F, Multigrid::Level const & C, int64_t c,
                         int64_t & parent, CellTransfer & T)
{
  int const nx = F.mesh.n_cells[0], ny = F.mesh.n_cells[1];
  int const cc[3] = {(int)(c % nx), (int)((c / nx) % ny), (int)(c / ((int64_t)nx * ny))};
  T.nf            = F.degree + 1;
  T.nc            = C.degree + 1;
  if(F.transfer == 'h')
  {
    int const pc[3] = {cc[0] / 2, cc[1] / 2, cc[2] / 2};
    parent = pc[0] + (int64_t)C.mesh.n_cells[0] * (pc[1] + (int64_t)C.mesh.n_cells[1] * pc[2]);
    for(int k = 0; k < 3; ++k)
      T.M[k] = F.Ph[cc[k] & 1].data();
  }
  else
  {
    parent = c;
    for(int k = 0; k < 3; ++k)
      T.M[k] = F.P.data();
  }
}
}  // namespace

void Multigrid::prolongate_add(int l, double * fine, double const * coarse) const
{
  Level const & F  = levels[l];
  Level const & C  = levels[l - 1];
  int64_t const nc = F.mesh.n_cells[0] * (int64_t)F.mesh.n_cells[1] * F.mesh.n_cells[2];
  if(F.transfer == 'c')  // DG_p += continuous Q_p node values
  {
    int const n3 = (F.degree + 1) * (F.degree + 1) * (F.degree + 1);
    for(int64_t c = 0; c < nc; ++c)
      for(int i = 0; i < n3; ++i)
      {
        int64_t const g = C.space.node(c, i);
        if(!C.space.constrained[g])
          fine[c * n3 + i] += coarse[g];
      }
    return;
  }
  int const nf3 = (F.degree + 1) * (F.degree + 1) * (F.degree + 1);
  int const nc3 = (C.degree + 1) * (C.degree + 1) * (C.degree + 1);
  if(F.dg)
  {
#pragma omp parallel for schedule(static)
    for(int64_t c = 0; c < nc; ++c)
    {
      int64_t      parent;
      CellTransfer T;
      parent_and_matrices(F, C, c, parent, T);
      std::vector<double> out(nf3);
      tensor_apply(T.nf, T.nf, T.nf, T.nc, T.M[0], T.M[1], T.M[2], coarse + parent * nc3, out.data());
      for(int i = 0; i < nf3; ++i)
        fine[c * nf3 + i] += out[i];
    }
    return;
  }
  std::vector<double> value(F.n, 0.0);
  std::vector<double> in(nc3), out(nf3);
  for(int64_t c = 0; c < nc; ++c)
  {
    int64_t      parent;
    CellTransfer T;
    parent_and_matrices(F, C, c, parent, T);
    for(int i = 0; i < nc3; ++i)
    {
      int64_t const g = C.space.node(parent, i);
      in[i]           = C.space.constrained[g] ? 0.0 : coarse[g];
    }
    tensor_apply(T.nf, T.nf, T.nf, T.nc, T.M[0], T.M[1], T.M[2], in.data(), out.data());
    for(int i = 0; i < nf3; ++i)
      value[F.space.node(c, i)] = out[i];
  }
  for(int64_t g = 0; g < F.n; ++g)
    if(!F.space.constrained[g])
      fine[g] += value[g];
}

void Multigrid::restrict_add(int l, double * coarse, double const * fine) const
{
  Level const & F  = levels[l];
  Level const & C  = levels[l - 1];
  int64_t const nc = F.mesh.n_cells[0] * (int64_t)F.mesh.n_cells[1] * F.mesh.n_cells[2];
  if(F.transfer == 'c')  // continuous node += sum over the cells sharing it
  {
    int const n3 = (F.degree + 1) * (F.degree + 1) * (F.degree + 1);
    for(int64_t c = 0; c < nc; ++c)
      for(int i = 0; i < n3; ++i)
      {
        int64_t const g = C.space.node(c, i);
        if(!C.space.constrained[g])
          coarse[g] += fine[c * n3 + i];
      }
    return;
  }
  int const nf3 = (F.degree + 1) * (F.degree + 1) * (F.degree + 1);
  int const nc3 = (C.degree + 1) * (C.degree + 1) * (C.degree + 1);
  if(F.dg)
  {
    // serial: several fine cells of an h-transfer add into one parent
    std::vector<double> out(nc3);
    for(int64_t c = 0; c < nc; ++c)
    {
      int64_t      parent;
      CellTransfer T;
      parent_and_matrices(F, C, c, parent, T);
      std::fill(out.begin(), out.end(), 0.0);
      tensor_apply_T(T.nf, T.nf, T.nf, T.nc, T.M[0], T.M[1], T.M[2], fine + c * nf3, out.data());
      for(int i = 0; i < nc3; ++i)
        coarse[parent * nc3 + i] += out[i];
    }
    return;
  }
  std::vector<double> in(nf3), out(nc3), acc(C.n, 0.0);
  for(int64_t c = 0; c < nc; ++c)
  {
    int64_t      parent;
    CellTransfer T;
    parent_and_matrices(F, C, c, parent, T);
    for(int i = 0; i < nf3; ++i)
    {
      int64_t const g = F.space.node(c, i);
      in[i]           = F.space.constrained[g] ? 0.0 : fine[g] / F.space.valence(g);
    }
    std::fill(out.begin(), out.end(), 0.0);
    tensor_apply_T(T.nf, T.nf, T.nf, T.nc, T.M[0], T.M[1], T.M[2], in.data(), out.data());
    for(int i = 0; i < nc3; ++i)
      acc[C.space.node(parent, i)] += out[i];
  }
  for(int64_t g = 0; g < C.n; ++g)
    if(!C.space.constrained[g])
      coarse[g] += acc[g];
}

void Multigrid::v_cycle(int l) const
{
  Level const & L = levels[l];
  if(l == 0)
  {
    if(desc.coarse_solver == 1)
    {
      coarse_cheb.vmult(L.sol, L.defect);
      last_coarse_its = coarse_cheb.degree;
      return;
    }
    std::vector<double> r = L.defect;
    if(desc.singular)
      subtract_mean(r);
    double const r0  = std::sqrt(dotv(L.n, r.data(), r.data()));
    double const tol = std::max(1e-12, desc.coarse_rel_tol * r0);
    last_coarse_its  = pcg_record(L.n, L.A, jacobi(L.inv_diag), r.data(), L.sol.data(), tol,
                                  desc.coarse_max_iter, desc.singular, nullptr, nullptr);
    return;
  }
  L.smoother.vmult(L.sol, L.defect);
  L.A(L.t.data(), L.sol.data());
  for(int64_t i = 0; i < L.n; ++i)
    L.t[i] = L.defect[i] - L.t[i];
  restrict_add(l, levels[l - 1].defect.data(), L.t.data());
  v_cycle(l - 1);
  prolongate_add(l, L.sol.data(), levels[l - 1].sol.data());
  L.smoother.step(L.sol, L.defect);
}

void Multigrid::vmult(double * dst, double const * src) const
{
  int const top = (int)levels.size() - 1;
  for(int l = 0; l < top; ++l)
    std::fill(levels[l].defect.begin(), levels[l].defect.end(), 0.0);
  std::copy(src, src + levels[top].n, levels[top].defect.begin());
  v_cycle(top);
  std::copy(levels[top].sol.begin(), levels[top].sol.end(), dst);
}

}  // namespace oracle
