#include "laplace_operator.h"

#include "block_kernels.h"

#include <algorithm>
#include <tuple>
#include <map>
#include <cmath>
#include <type_traits>

namespace ExaDG
{
namespace Poisson
{
namespace
{
// The even-odd form of a centrosymmetric (skew = false) or skew-centrosymmetric (skew = true)
// n x n matrix (row-major) that the kernels read (CartTable, GenTable, eo_size); the plain
// matrix unless packed. Throws if A lacks the symmetry, which holds for the 1D mass and stiffness
// on symmetric nodes (centro), the Gauss interpolation S (centro) and the Gauss collocation
// derivative (skew).
std::vector<double> even_odd_pack(std::vector<double> const & A, int n, bool packed, bool skew)
{
  if(!packed)
    return A;
  double amax = 0;
  for(double v : A)
    amax = std::max(amax, std::abs(v));
  double const sign = skew ? -1.0 : 1.0;
  for(int i = 0; i < n; ++i)
    for(int j = 0; j < n; ++j)
      if(std::abs(A[i * n + j] - sign * A[(n - 1 - i) * n + (n - 1 - j)]) > 1e-12 * amax)
        EXADG_THROW(EXADG_ERR_INTERNAL, skew ? "1D matrix is not skew-centrosymmetric"
                                             : "1D matrix is not centrosymmetric");
  int const           m = n / 2, me = m + n % 2;
  std::vector<double> t;
  for(int i = 0; i < m; ++i)
    for(int j = 0; j < me; ++j)
      t.push_back(j < m ? 0.5 * (A[i * n + j] + A[i * n + n - 1 - j]) : A[i * n + m]);
  for(int i = 0; i < m; ++i)
    for(int j = 0; j < m; ++j)
      t.push_back(0.5 * (A[i * n + j] - A[i * n + n - 1 - j]));
  if(n % 2)
    for(int j = 0; j <= m; ++j)  // middle row: coefficients of e_j (centro) or o_j (skew)
      t.push_back(skew ? (j < m ? A[m * n + j] : 0.0) : A[m * n + j]);
  if((int)t.size() != eo_size(n, true))
    EXADG_THROW(EXADG_ERR_INTERNAL, "even-odd table size");
  return t;
}

std::vector<double> transposed(std::vector<double> const & A, int n)
{
  std::vector<double> t((size_t)n * n);
  for(int i = 0; i < n; ++i)
    for(int j = 0; j < n; ++j)
      t[j * n + i] = A[i * n + j];
  return t;
}
}  // namespace

template<typename Number>
LaplaceOperator<Number>::LaplaceOperator(std::shared_ptr<StructuredGrid const> grid,
                                         LaplaceOperatorData const &           data)
  : grid_(std::move(grid)), data_(data)
{
  if(data.degree < 1 || data.degree > 8)
    EXADG_THROW(EXADG_ERR_ARGUMENT, "degree must be in [1,8]");
  shape_     = make_shape_info(data.degree);
  int const n = shape_.n;
  n3_         = n * n * n;
  bool const curved = !grid_->affine();
  geo_              = compute_geometry(*grid_, shape_, data.IP_factor, curved);

  // 1D tables travel by value in the kernel arguments (uniform scalar loads); the general
  // kernel reads the plain layout from device memory
  std::vector<double> tab, plain;
  auto append = [](std::vector<double> & t, std::vector<double> const & v) {
    t.insert(t.end(), v.begin(), v.end());
  };
  if(curved)
  {
    append(tab, shape_.S);
    append(tab, shape_.Dq);
  }
  else
  {
    append(tab, even_odd_pack(shape_.mass, n, cs_packed(n), false));
    append(tab, even_odd_pack(shape_.stiff, n, cs_packed(n), false));
  }
  append(tab, shape_.d0);
  append(tab, shape_.d1);
  append(plain, curved ? shape_.S : shape_.mass);
  append(plain, curved ? shape_.Dq : shape_.stiff);
  append(plain, shape_.d0);
  append(plain, shape_.d1);
  if(curved)
  {
    // Lagrange polynomials on the Gauss points at the cell ends: values l_q(0), l_q(1), then
    // derivatives l_q'(0), l_q'(1) (the general kernel's face traces from Gauss values)
    std::vector<double> v0(n), v1(n), g0(n), g1(n);
    lagrange_basis(shape_.qp, 0.0, v0.data(), g0.data());
    lagrange_basis(shape_.qp, 1.0, v1.data(), g1.data());
    append(plain, v0);
    append(plain, v1);
    append(plain, g0);
    append(plain, g1);
    // the sweep matrices S, Dq, S^T, Dq^T in even-odd form (GenTable)
    bool const gp = gen_packed(n);
    append(plain, even_odd_pack(shape_.S, n, gp, false));
    append(plain, even_odd_pack(shape_.Dq, n, gp, true));
    append(plain, even_odd_pack(transposed(shape_.S, n), n, gp, false));
    append(plain, even_odd_pack(transposed(shape_.Dq, n), n, gp, true));
    int const size = 2 * n * n + 6 * n + 4 * gen_size(n);
    if((int)plain.size() != size)
      EXADG_THROW(EXADG_ERR_INTERNAL, "general-kernel table size");
  }
  if((int)tab.size() > SipgArgs<Number>::kMaxTable)
    EXADG_THROW(EXADG_ERR_INTERNAL, "1D tables exceed the kernel-argument table");
  tab_.assign(tab.begin(), tab.end());
  tables_.upload(std::vector<Number>(plain.begin(), plain.end()));
  if(!curved)
  {
    // unit-interval 1D diagonals and end derivatives of the closed-form diagonal
    std::vector<double> dd(4 * (size_t)n);
    for(int i = 0; i < n; ++i)
    {
      dd[i]         = shape_.mass[i * n + i];
      dd[n + i]     = shape_.stiff[i * n + i];
      dd[2 * n + i] = shape_.d0[i];
      dd[3 * n + i] = shape_.d1[i];
    }
    diag1d_.upload(dd);
  }
  else
  {
    // tables of the sum-factorized curved diagonal: products of the Gauss values S and the
    // derivatives D = Dq S of the nodal basis, [q][i] each, then phi_0'(0) and phi_{n-1}'(1)
    std::vector<double> D((size_t)n * n, 0.0), t((size_t)3 * n * n + 2);
    for(int q = 0; q < n; ++q)
      for(int i = 0; i < n; ++i)
        for(int m = 0; m < n; ++m)
          D[q * n + i] += shape_.Dq[q * n + m] * shape_.S[m * n + i];
    for(int k = 0; k < n * n; ++k)
    {
      t[k]             = shape_.S[k] * shape_.S[k];
      t[n * n + k]     = D[k] * D[k];
      t[2 * n * n + k] = shape_.S[k] * D[k];
    }
    t[3 * n * n]     = shape_.d0[0];
    t[3 * n * n + 1] = shape_.d1[n - 1];
    diag_tab_.upload(t);
  }
  for(int d = 0; d < 3; ++d)
  {
    h_[d]     = (Number)grid_->h(d);
    h_[3 + d] = (Number)(1.0 / grid_->h(d));
  }
  {
    // FP64 tables of the divergence right-hand side: S, Sd = Dq S, then weights and points
    std::vector<double> dt(shape_.S);
    std::vector<double> Sd((size_t)n * n, 0.0);
    for(int q = 0; q < n; ++q)
      for(int i = 0; i < n; ++i)
        for(int p = 0; p < n; ++p)
          Sd[q * n + i] += shape_.Dq[q * n + p] * shape_.S[p * n + i];
    dt.insert(dt.end(), Sd.begin(), Sd.end());
    dt.insert(dt.end(), shape_.qw.begin(), shape_.qw.end());
    dt.insert(dt.end(), shape_.qp.begin(), shape_.qp.end());
    div_tables_.upload(dt);
    // right-hand side tables: S, Sd, d0, d1
    std::vector<double> rt(shape_.S);
    rt.insert(rt.end(), Sd.begin(), Sd.end());
    rt.insert(rt.end(), shape_.d0.begin(), shape_.d0.end());
    rt.insert(rt.end(), shape_.d1.begin(), shape_.d1.end());
    rhs_tables_.upload(rt);
  }
  {
    // boundary faces of the non-periodic sides, side-major (the RhsGeometry numbering)
    BoxTopology const & t = grid_->topology();
    for(int side = 0; side < 6; ++side)
    {
      int const d = side / 2;
      rhs_side_base_[side + 1] =
        rhs_side_base_[side] + (t.periodic[d] ? 0 : (int64_t)t.nc[tangent1(d)] * t.nc[tangent2(d)]);
    }
  }
  {
    // error norm with QGauss(k+3): Se[q][i] = phi_i(e_q), then weights, then points
    int const           ne = n + 2;
    std::vector<double> ep, ew, Se((size_t)ne * n), val(n), der(n);
    gauss_points(ne, ep, ew);
    for(int q = 0; q < ne; ++q)
    {
      lagrange_basis(shape_.nodes, ep[q], val.data(), der.data());
      for(int i = 0; i < n; ++i)
        Se[q * n + i] = val[i];
    }
    Se.insert(Se.end(), ew.begin(), ew.end());
    Se.insert(Se.end(), ep.begin(), ep.end());
    err_tables_.upload(Se);
  }
  auto convert = [](std::vector<double> const & v) {
    return std::vector<Number>(v.begin(), v.end());
  };
  tau_face_.upload(convert(geo_.tau_face));
  if(curved)
  {
    cell_metric_.upload(convert(geo_.cell_metric));
    face_data_.upload(convert(geo_.face_data));
    // the metric lives on the device; keep only tau on the host
    geo_.cell_metric.clear();
    geo_.cell_metric.shrink_to_fit();
    geo_.face_data.clear();
    geo_.face_data.shrink_to_fit();
  }
}

template<typename Number>
SipgArgs<Number> LaplaceOperator<Number>::make_args(Number * dst, Number const * src, int probe) const
{
  SipgArgs<Number> a;
  std::copy(tab_.begin(), tab_.end(), a.tab);
  std::copy(h_, h_ + 6, a.h);
  a.tables      = tables_.data();
  a.src         = src;
  a.dst         = dst;
  a.cell_metric = cell_metric_.data();
  a.face_data   = face_data_.data();
  a.tau_face    = tau_face_.data();
  a.topo        = grid_->topology();
  a.n_cells     = grid_->n_cells();
  a.probe       = probe;
  // one tile of y-rows per XCD
  a.tile_rows = std::max(1, (grid_->topology().nc[1] + 7) / 8);
  return a;
}

template<typename Number>
void LaplaceOperator<Number>::launch(ApplyMode mode, SipgArgs<Number> const & a, hipStream_t s) const
{
  hipError_t const e = grid_->affine() ? launch_sipg_cartesian<Number>(data_.degree, mode, a, s)
                                       : launch_sipg_general<Number>(data_.degree, mode, a, s);
  if(e != hipSuccess)
    EXADG_THROW(EXADG_ERR_HIP, std::string("SIPG kernel launch: ") + hipGetErrorString(e));
}

template<typename Number>
void LaplaceOperator<Number>::apply(ApplyMode mode, Number * dst, Number const * src, int probe,
                                    hipStream_t s) const
{
  if(mode != ApplyMode::diagonal && (dst == nullptr || src == nullptr))
    EXADG_THROW(EXADG_ERR_ARGUMENT, "null vector");
  if(mode != ApplyMode::diagonal && (void const *)dst == (void const *)src)
    EXADG_THROW(EXADG_ERR_ARGUMENT, "dst and src must not alias");
  launch(mode, make_args(dst, src, probe), s);
}

template<typename Number>
void LaplaceOperator<Number>::chebyshev_step(Number * x_new, Number const * x, Number const * x_old,
                                             Number const * b, Number const * dinv, Number f1,
                                             Number f2, bool has_old, hipStream_t s) const
{
  if(x_new == nullptr || x == nullptr || b == nullptr || dinv == nullptr || (has_old && !x_old))
    EXADG_THROW(EXADG_ERR_ARGUMENT, "null vector");
  if((void const *)x_new == (void const *)x || (void const *)b == (void const *)x_new ||
     (void const *)dinv == (void const *)x_new)
    EXADG_THROW(EXADG_ERR_ARGUMENT, "x_new must not alias x, b or dinv");
  SipgArgs<Number> a = make_args(x_new, x, 0);
  a.x_old            = x_old;
  a.rhs              = b;
  a.dinv             = dinv;
  a.f1               = f1;
  a.f2               = f2;
  a.has_old          = has_old ? 1 : 0;
  launch(ApplyMode::chebyshev, a, s);
}

template<typename Number>
void LaplaceOperator<Number>::residual(Number * dst, Number const * src, Number const * b,
                                       hipStream_t s) const
{
  if(dst == nullptr || src == nullptr || b == nullptr)
    EXADG_THROW(EXADG_ERR_ARGUMENT, "null vector");
  if((void const *)dst == (void const *)src || (void const *)b == (void const *)src)
    EXADG_THROW(EXADG_ERR_ARGUMENT, "dst and b must not alias src");
  SipgArgs<Number> a = make_args(dst, src, 0);
  a.rhs              = b;
  launch(ApplyMode::residual, a, s);
}

template<typename Number>
bool LaplaceOperator<Number>::residual_restrict(Number * coarse, Number const * src, Number const * b,
                                                Number const * P, int nc, hipStream_t s) const
{
  if(coarse == nullptr || src == nullptr || b == nullptr || P == nullptr)
    EXADG_THROW(EXADG_ERR_ARGUMENT, "null vector");
  if(nc < 1 || nc > data_.degree + 1)
    EXADG_THROW(EXADG_ERR_ARGUMENT, "coarse degree out of range");
  SipgArgs<Number> a = make_args(coarse, src, 0);
  a.rhs              = b;
  a.restrict_P       = P;
  a.restrict_nc      = nc;
  launch(ApplyMode::residual_restrict, a, s);
  return true;
}

template<typename Number>
void LaplaceOperator<Number>::vmult(Number * dst, Number const * src, hipStream_t s) const
{
  apply(ApplyMode::vmult, dst, src, 0, s);
}

template<typename Number>
void LaplaceOperator<Number>::vmult_add(Number * dst, Number const * src, hipStream_t s) const
{
  apply(ApplyMode::vmult_add, dst, src, 0, s);
}

template<typename Number>
void LaplaceOperator<Number>::calculate_diagonal(Number * diag, hipStream_t s) const
{
  if(diag == nullptr)
    EXADG_THROW(EXADG_ERR_ARGUMENT, "null vector");
  if(grid_->affine())
  {
    // closed form: one launch (the Kronecker-sum cell block's diagonal)
    double const     h[3] = {grid_->h(0), grid_->h(1), grid_->h(2)};
    hipError_t const e    = launch_cartesian_diagonal<Number>(diag, diag1d_.data(), tau_face_.data(),
                                                           grid_->topology(), h, shape_.n,
                                                           grid_->n_cells(), s);
    if(e != hipSuccess)
      EXADG_THROW(EXADG_ERR_HIP, hipGetErrorString(e));
    return;
  }
  BoxTopology const & t          = grid_->topology();
  bool                self_periodic = false;
  for(int d = 0; d < 3; ++d)
    self_periodic = self_periodic || (t.periodic[d] && t.nc[d] == 1);
  if(!self_periodic)
  {
    // curved: sum-factorized cell and face terms, one launch
    hipError_t const e = launch_general_diagonal<Number>(diag, diag_tab_.data(),
                                                         make_args(diag, nullptr, 0), data_.degree, s);
    if(e != hipSuccess)
      EXADG_THROW(EXADG_ERR_HIP, hipGetErrorString(e));
    return;
  }
  // a cell that is its own neighbour: one launch per local unit vector, each writing entry
  // `probe` of every cell
  for(int probe = 0; probe < n3_; ++probe)
    apply(ApplyMode::diagonal, diag, nullptr, probe, s);
}

template<typename Number>
void LaplaceOperator<Number>::calculate_inverse_diagonal(Number * diag, hipStream_t s) const
{
  calculate_diagonal(diag, s);
  invert_diagonal_device<Number>(diag, m(), s);
}

namespace
{
// eigenvalues and eigenvectors of a symmetric n x n matrix (cyclic Jacobi rotations); on return
// A holds the eigenvalues on its diagonal and Q (row-major) the eigenvectors as columns
void symmetric_eigen(std::vector<double> & A, std::vector<double> & Q, int n)
{
  Q.assign((size_t)n * n, 0.0);
  for(int i = 0; i < n; ++i)
    Q[i * n + i] = 1.0;
  for(int sweep = 0; sweep < 100; ++sweep)
  {
    double off = 0, diag = 0;
    for(int i = 0; i < n; ++i)
      for(int j = 0; j < n; ++j)
        (i == j ? diag : off) += A[i * n + j] * A[i * n + j];
    if(off <= 1e-30 * diag)
      return;
    for(int p = 0; p < n; ++p)
      for(int q = p + 1; q < n; ++q)
      {
        double const apq = A[p * n + q];
        if(apq == 0.0)
          continue;
        double const theta = 0.5 * (A[q * n + q] - A[p * n + p]) / apq;
        double const t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        double const c = 1.0 / std::sqrt(t * t + 1.0), sn = t * c;
        for(int k = 0; k < n; ++k)  // A <- A R (columns p, q)
        {
          double const akp = A[k * n + p], akq = A[k * n + q];
          A[k * n + p]     = c * akp - sn * akq;
          A[k * n + q]     = sn * akp + c * akq;
        }
        for(int k = 0; k < n; ++k)  // A <- R^T A (rows p, q)
        {
          double const apk = A[p * n + k], aqk = A[q * n + k];
          A[p * n + k]     = c * apk - sn * aqk;
          A[q * n + k]     = sn * apk + c * aqk;
        }
        for(int k = 0; k < n; ++k)
        {
          double const qkp = Q[k * n + p], qkq = Q[k * n + q];
          Q[k * n + p]     = c * qkp - sn * qkq;
          Q[k * n + q]     = sn * qkp + c * qkq;
        }
      }
  }
  EXADG_THROW(EXADG_ERR_INTERNAL, "Jacobi eigenvalue iteration did not converge");
}

// the 1D cell-block operator along one direction (the line Laplacian of the affine kernel with
// zero neighbour values, or the cell's own values across a periodic face of a one-cell
// direction), and its generalized eigen-decomposition with the mass h M
void cell_line_eigen(ShapeInfo1D const & sh, double h, int type_l, int type_r, bool self, double tau_l,
                     double tau_r, double * vec, double * val)
{
  int const           n = sh.n;
  std::vector<double> L((size_t)n * n), u(n), o(n);
  for(int j = 0; j < n; ++j)
  {
    std::fill(u.begin(), u.end(), 0.0);
    u[j] = 1.0;
    for(int i = 0; i < n; ++i)
    {
      o[i] = 0;
      for(int k = 0; k < n; ++k)
        o[i] += sh.stiff[i * n + k] * u[k];
    }
    double du0 = 0, du1 = 0;
    for(int i = 0; i < n; ++i)
    {
      du0 += sh.d0[i] * u[i];
      du1 += sh.d1[i] * u[i];
    }
    auto exterior = [&](int type, bool left, double & w, double & dw) {
      double const ue = left ? u[0] : u[n - 1], de = left ? du0 : du1;
      if(type == EXADG_BC_DIRICHLET)
        w = -ue, dw = de;
      else if(type == EXADG_BC_NEUMANN)
        w = ue, dw = -de;
      else if(self)
        w = left ? u[n - 1] : u[0], dw = left ? du1 : du0;
      else
        w = 0, dw = 0;
    };
    double wl, dwl, wr, dwr;
    exterior(type_l, true, wl, dwl);
    exterior(type_r, false, wr, dwr);
    double const jl = 0.5 * (u[0] - wl), jr = 0.5 * (u[n - 1] - wr);
    for(int i = 0; i < n; ++i)
      o[i] = (o[i] + sh.d0[i] * jl - sh.d1[i] * jr) / h;
    o[0] += 0.5 * (du0 + dwl) / h + 2.0 * tau_l * jl;
    o[n - 1] += 2.0 * tau_r * jr - 0.5 * (du1 + dwr) / h;
    for(int i = 0; i < n; ++i)
      L[i * n + j] = o[i];
  }
  double asym = 0, scale = 0;
  for(int i = 0; i < n; ++i)
    for(int j = 0; j < n; ++j)
    {
      asym  = std::max(asym, std::fabs(L[i * n + j] - L[j * n + i]));
      scale = std::max(scale, std::fabs(L[i * n + j]));
    }
  if(asym > 1e-10 * scale)
    EXADG_THROW(EXADG_ERR_INTERNAL, "1D cell block is not symmetric");
  // h M = C C^T (Cholesky), B = C^{-1} L C^{-T}, B = Q diag Q^T, V = C^{-T} Q
  std::vector<double> C((size_t)n * n, 0.0);
  for(int j = 0; j < n; ++j)
  {
    double d = h * sh.mass[j * n + j];
    for(int k = 0; k < j; ++k)
      d -= C[j * n + k] * C[j * n + k];
    C[j * n + j] = std::sqrt(d);
    for(int i = j + 1; i < n; ++i)
    {
      double v = h * sh.mass[i * n + j];
      for(int k = 0; k < j; ++k)
        v -= C[i * n + k] * C[j * n + k];
      C[i * n + j] = v / C[j * n + j];
    }
  }
  std::vector<double> Y((size_t)n * n), B((size_t)n * n), Q;
  for(int col = 0; col < n; ++col)  // Y = C^{-1} L (forward substitution per column)
    for(int i = 0; i < n; ++i)
    {
      double v = 0.5 * (L[i * n + col] + L[col * n + i]);
      for(int k = 0; k < i; ++k)
        v -= C[i * n + k] * Y[k * n + col];
      Y[i * n + col] = v / C[i * n + i];
    }
  for(int row = 0; row < n; ++row)  // B = Y C^{-T}: solve C B^T = Y^T row by row
    for(int i = 0; i < n; ++i)
    {
      double v = Y[row * n + i];
      for(int k = 0; k < i; ++k)
        v -= C[i * n + k] * B[row * n + k];
      B[row * n + i] = v / C[i * n + i];
    }
  for(int i = 0; i < n; ++i)  // symmetrize the rounding
    for(int j = i + 1; j < n; ++j)
      B[i * n + j] = B[j * n + i] = 0.5 * (B[i * n + j] + B[j * n + i]);
  symmetric_eigen(B, Q, n);
  // positive semidefinite: a one-cell periodic direction has the constant mode (zero up to
  // rounding); the sum over the three directions must stay positive (checked per cell)
  double vmax = 0;
  for(int a = 0; a < n; ++a)
    vmax = std::max(vmax, std::fabs(B[a * n + a]));
  for(int a = 0; a < n; ++a)
  {
    val[a] = B[a * n + a];
    if(val[a] < -1e-12 * vmax)
      EXADG_THROW(EXADG_ERR_INTERNAL, "1D cell block is not positive semidefinite");
    val[a] = std::max(val[a], 0.0);
  }
  for(int col = 0; col < n; ++col)  // V = C^{-T} Q: back substitution per column
    for(int i = n - 1; i >= 0; --i)
    {
      double v = Q[i * n + col];
      for(int k = i + 1; k < n; ++k)
        v -= C[k * n + i] * vec[k * n + col];
      vec[i * n + col] = v / C[i * n + i];
    }
}
}  // namespace

template<typename Number>
void LaplaceOperator<Number>::fast_diagonalization_tables(std::vector<double> & vec,
                                                          std::vector<double> & val,
                                                          std::vector<int> &    cfg) const
{
  if(!grid_->affine())
    EXADG_THROW(EXADG_ERR_ARGUMENT, "fast diagonalization needs an affine mesh");
  BoxTopology const & t  = grid_->topology();
  int const           n  = shape_.n;
  int64_t const       nc = grid_->n_cells();
  using Key              = std::tuple<int, int, int, int, double, double>;
  std::map<Key, int>  index;
  vec.clear();
  val.clear();
  cfg.assign((size_t)nc * 3, 0);
  for(int64_t c = 0; c < nc; ++c)
  {
    int cc[3];
    cell_coords(t, c, cc);
    for(int d = 0; d < 3; ++d)
    {
      int64_t const nl = neighbor_id(t, cc, d, 0), nr = neighbor_id(t, cc, d, 1);
      int const     tl = nl >= 0 ? -1 : t.bc[2 * d], tr = nr >= 0 ? -1 : t.bc[2 * d + 1];
      bool const    self = nl == c;
      double const  taul = geo_.tau_face[face_id(t, cc, d, 0)];
      double const  taur = geo_.tau_face[face_id(t, cc, d, 1)];
      Key const     key{d, tl, tr, self ? 1 : 0, taul, taur};
      auto          it = index.find(key);
      if(it == index.end())
      {
        int const id = (int)index.size();
        it           = index.emplace(key, id).first;
        vec.resize((size_t)(id + 1) * n * n);
        val.resize((size_t)(id + 1) * n);
        cell_line_eigen(shape_, grid_->h(d), tl, tr, self, taul, taur, &vec[(size_t)id * n * n],
                        &val[(size_t)id * n]);
      }
      cfg[(size_t)c * 3 + d] = it->second;
    }
    double lmin = 0, lmax = 0;
    for(int d = 0; d < 3; ++d)
    {
      double const * v = &val[(size_t)cfg[(size_t)c * 3 + d] * n];
      lmin += *std::min_element(v, v + n);
      lmax += *std::max_element(v, v + n);
    }
    if(!(lmin > 1e-12 * lmax))
      EXADG_THROW(EXADG_ERR_ARGUMENT, "a cell block is singular (a one-cell box periodic in "
                                      "every direction); block Jacobi is undefined");
  }
}

template<typename Number>
CellBlockInverse<Number>::CellBlockInverse(LaplaceOperator<Number> const & op, hipStream_t s)
  : n_cells_(op.n_cells())
{
  m_   = (int)(op.m() / op.n_cells());
  n1d_ = op.degree() + 1;
  if(op.is_affine())
  {
    fdm_ = true;
    std::vector<double> vec, val;
    std::vector<int>    cfg;
    op.fast_diagonalization_tables(vec, val, cfg);
    vec_.upload(std::vector<Number>(vec.begin(), vec.end()));
    val_.upload(std::vector<Number>(val.begin(), val.end()));
    cfg_.upload(cfg);
    return;
  }
  if(m_ > kMaxDenseBlock)
    EXADG_THROW(EXADG_ERR_ARGUMENT,
                "block Jacobi on curved meshes holds blocks up to 125 x 125 (degree <= 4)");
  blocks_.resize((size_t)n_cells_ * m_ * m_);
  op.calculate_inverse_block_diagonal(blocks_.data(), s);
}

template<typename Number>
void CellBlockInverse<Number>::apply(BlockApplyArgs<Number> a, hipStream_t s) const
{
  a.m       = m_;
  a.n_cells = n_cells_;
  if(fdm_)
  {
    a.fdm_vec = vec_.data();
    a.fdm_val = val_.data();
    a.fdm_cfg = cfg_.data();
    a.n1d     = n1d_;
  }
  else
    a.blocks = blocks_.data();
  hipError_t e = launch_block_apply<Number>(a, s);
  if(e != hipSuccess)
    EXADG_THROW(EXADG_ERR_HIP, hipGetErrorString(e));
}

template<typename Number>
void LaplaceOperator<Number>::calculate_block_diagonal(Number * blocks, hipStream_t s) const
{
  if(blocks == nullptr)
    EXADG_THROW(EXADG_ERR_ARGUMENT, "null vector");
  for(int probe = 0; probe < n3_; ++probe)
  {
    SipgArgs<Number> a = make_args(blocks, nullptr, probe);
    a.block            = 1;
    launch(ApplyMode::diagonal, a, s);
  }
}

template<typename Number>
void LaplaceOperator<Number>::calculate_inverse_block_diagonal(Number * blocks, hipStream_t s) const
{
  if(n3_ > kMaxDenseBlock)
    EXADG_THROW(EXADG_ERR_ARGUMENT, "block Jacobi holds blocks up to 125 x 125 (degree <= 4)");
  calculate_block_diagonal(blocks, s);
  hipError_t e = launch_block_invert<Number>(blocks, n3_, n_cells(), s);
  if(e != hipSuccess)
    EXADG_THROW(EXADG_ERR_HIP, hipGetErrorString(e));
}

template<typename Number>
void LaplaceOperator<Number>::ensure_rhs_geometry() const
{
  if(rhs_cell_geo_.size() > 0)
    return;
  RhsGeometry const g = compute_rhs_geometry(*grid_, shape_, geo_.tau_face);
  rhs_cell_geo_.upload(g.cell);
  if(!g.face.empty())
    rhs_face_geo_.upload(g.face);
}

template<typename Number>
void LaplaceOperator<Number>::rhs_launch(double const * f, double const * boundary, int sine,
                                         double omega, double * b, hipStream_t s) const
{
  if(b == nullptr)
    EXADG_THROW(EXADG_ERR_ARGUMENT, "null vector");
  ensure_rhs_geometry();
  RhsArgs a;
  a.b        = b;
  a.cell_geo = rhs_cell_geo_.data();
  a.face_geo = rhs_face_geo_.data();
  a.f        = f;
  a.boundary = boundary;
  a.tables   = rhs_tables_.data();
  a.topo     = grid_->topology();
  std::copy(rhs_side_base_, rhs_side_base_ + 7, a.side_base);
  a.n       = shape_.n;
  a.sine    = sine;
  a.omega   = omega;
  a.n_cells = grid_->n_cells();
  hipError_t const e = launch_rhs_assemble(a, s);
  if(e != hipSuccess)
    EXADG_THROW(EXADG_ERR_HIP, std::string("rhs kernel launch: ") + hipGetErrorString(e));
}

template<typename Number>
void LaplaceOperator<Number>::rhs_sine(double omega, double * b, hipStream_t s) const
{
  rhs_launch(nullptr, nullptr, 1, omega, b, s);
}

template<typename Number>
void LaplaceOperator<Number>::rhs(double const * f, double const * boundary, double * b,
                                  hipStream_t s) const
{
  if((f != nullptr && (void const *)f == (void const *)b) ||
     (boundary != nullptr && (void const *)boundary == (void const *)b))
    EXADG_THROW(EXADG_ERR_ARGUMENT, "b must not alias the point values");
  rhs_launch(f, boundary, 0, 0.0, b, s);
}

template<typename Number>
void LaplaceOperator<Number>::rhs_n_points(int64_t & n_cell_points, int64_t & n_boundary_points) const
{
  int const n     = shape_.n;
  n_cell_points     = grid_->n_cells() * n3_;
  n_boundary_points = rhs_side_base_[6] * n * n;
}

template<typename Number>
void LaplaceOperator<Number>::rhs_points(double * x_cell, double * x_boundary, hipStream_t s) const
{
  ensure_rhs_geometry();
  int const     n = shape_.n, n2 = n * n;
  int64_t const nc = grid_->n_cells(), nb = rhs_side_base_[6];
  // records are SoA per cell / face; the output is component-major over all points
  for(int d = 0; d < 3; ++d)
  {
    if(x_cell)
      HIP_CHECK(hipMemcpy2DAsync(x_cell + d * nc * n3_, n3_ * sizeof(double),
                                 rhs_cell_geo_.data() + d * n3_, kRhsCellRec * n3_ * sizeof(double),
                                 n3_ * sizeof(double), nc, hipMemcpyDeviceToDevice, s));
    if(x_boundary && nb > 0)
      HIP_CHECK(hipMemcpy2DAsync(x_boundary + d * nb * n2, n2 * sizeof(double),
                                 rhs_face_geo_.data() + d * n2, kRhsFaceRec * n2 * sizeof(double),
                                 n2 * sizeof(double), nb, hipMemcpyDeviceToDevice, s));
  }
}

template<typename Number>
bool LaplaceOperator<Number>::is_singular() const
{
  BoxTopology const & t = grid_->topology();
  for(int d = 0; d < 3; ++d)
    for(int s = 0; s < 2; ++s)
      if(!t.periodic[d] && t.bc[2 * d + s] == EXADG_BC_DIRICHLET)
        return false;
  return true;
}

template<typename Number>
void LaplaceOperator<Number>::rhs_divergence_fourier(double const * modes_host, int n_modes,
                                                     double t, double * b, hipStream_t s) const
{
  if(b == nullptr || modes_host == nullptr || n_modes < 1)
    EXADG_THROW(EXADG_ERR_ARGUMENT, "invalid Fourier field");
  // per-call mode data: n, d, and c = A exp(i (phi + omega t)), so that the kernel evaluates
  // u = sum A d cos(2 pi n.x + phi + omega t) = Re sum d c exp(i 2 pi n.x)
  std::vector<double> mh((size_t)8 * n_modes);
  int                 K = 0;
  for(int m = 0; m < n_modes; ++m)
  {
    double const * q = modes_host + 10 * m;
    for(int d = 0; d < 3; ++d)
    {
      if(q[d] != std::round(q[d]) || std::abs(q[d]) > 16)
        EXADG_THROW(EXADG_ERR_ARGUMENT, "wave vectors must be integer with |n_d| <= 16");
      K = std::max(K, (int)std::abs(q[d]));
      mh[8 * m + d]     = q[d];
      mh[8 * m + 3 + d] = q[3 + d];
    }
    double const ph = q[7] + q[8] * t;
    mh[8 * m + 6]   = q[6] * std::cos(ph);
    mh[8 * m + 7]   = q[6] * std::sin(ph);
  }
  if(modes_.size() != mh.size())
    modes_.resize(mh.size());
  HIP_CHECK(hipMemcpyAsync(modes_.data(), mh.data(), mh.size() * sizeof(double), hipMemcpyHostToDevice, s));
  int const    n    = shape_.n;
  double const h[3] = {grid_->h(0), grid_->h(1), grid_->h(2)};
  hipError_t   e    = hipSuccess;
  if(grid_->affine())
    e = launch_divergence_rhs(b, modes_.data(), n_modes, K, grid_->topology(), grid_->mesh().lower,
                              h, n, div_tables_.data(), div_tables_.data() + n * n,
                              div_tables_.data() + 2 * n * n, grid_->n_cells(), s);
  else
  {
    if(div_cell_geo_.size() == 0)  // point data of the curved mesh, computed on first use
    {
      DivergenceGeometry const g = compute_divergence_geometry(*grid_, shape_);
      div_cell_geo_.upload(g.cell);
      div_face_geo_.upload(g.face);
    }
    e = launch_divergence_rhs_curved(b, modes_.data(), n_modes, K, grid_->topology(), n,
                                     div_tables_.data(), div_tables_.data() + n * n,
                                     div_tables_.data() + 2 * n * n, div_cell_geo_.data(),
                                     div_face_geo_.data(), grid_->n_cells(), s);
  }
  if(e != hipSuccess)
    EXADG_THROW(EXADG_ERR_HIP, hipGetErrorString(e));
  if(is_singular())
  {
    if(!reducer_)
      reducer_ = std::make_unique<Reducer>();
    vec::subtract_mean<double>(*reducer_, b, m(), s);
  }
  HIP_CHECK(hipStreamSynchronize(s));
}

template<typename Number>
double LaplaceOperator<Number>::error_sine(double omega, double const * u, hipStream_t s) const
{
  if(!std::is_same<Number, double>::value)
    EXADG_THROW(EXADG_ERR_ARGUMENT, "the error norm needs the FP64 operator");
  if(u == nullptr)
    EXADG_THROW(EXADG_ERR_ARGUMENT, "null vector");
  int const     n = shape_.n, ne = n + 2, ne3 = ne * ne * ne;
  int64_t const nc = grid_->n_cells();
  double const * Se = err_tables_.data();
  double const * w  = Se + ne * n;
  if(!grid_->affine() && err_geo_.size() == 0)
  {
    // positions and JxW at the error points of every cell, computed once
    std::vector<double> hw(2 * ne);
    HIP_CHECK(hipMemcpy(hw.data(), w, sizeof(double) * 2 * ne, hipMemcpyDeviceToHost));
    std::vector<double> ew(hw.begin(), hw.begin() + ne), epts(hw.begin() + ne, hw.end());
    std::vector<double> geo((size_t)nc * 4 * ne3);
    parallel_range(nc, [&](int64_t b, int64_t e) {
      std::vector<double> x(3 * ne3), J(9 * ne3);
      for(int64_t c = b; c < e; ++c)
      {
        grid_->evaluate(c, epts, epts, epts, x.data(), J.data());
        double * g = &geo[(size_t)c * 4 * ne3];
        for(int qz = 0; qz < ne; ++qz)
          for(int qy = 0; qy < ne; ++qy)
            for(int qx = 0; qx < ne; ++qx)
            {
              int const q = qx + ne * (qy + ne * qz);
              double    a[9];
              for(int k = 0; k < 9; ++k)
                a[k] = J[k * ne3 + q];
              double const det = a[0] * (a[4] * a[8] - a[5] * a[7]) -
                                 a[1] * (a[3] * a[8] - a[5] * a[6]) +
                                 a[2] * (a[3] * a[7] - a[4] * a[6]);
              for(int d = 0; d < 3; ++d)
                g[d * ne3 + q] = x[d * ne3 + q];
              g[3 * ne3 + q] = std::abs(det) * ew[qx] * ew[qy] * ew[qz];
            }
      }
    });
    err_geo_.upload(geo);
  }
  if(err_partial_.size() < (size_t)(2 * nc))
    err_partial_.resize((size_t)(2 * nc));
  if(err_scratch_.size() < 2 * 1024 + 2)
    err_scratch_.resize(2 * 1024 + 2);
  double const h[3] = {grid_->h(0), grid_->h(1), grid_->h(2)};
  hipError_t   e    = launch_error_sine(err_partial_.data(), u, omega, n, ne, Se, w,
                                        grid_->topology(), grid_->mesh().lower, h,
                                        grid_->affine() ? nullptr : err_geo_.data(), nc, s);
  if(e == hipSuccess)
    e = launch_sum_pairs(err_partial_.data(), nc, err_scratch_.data(),
                         err_scratch_.data() + 2 * 1024, s);
  if(e != hipSuccess)
    EXADG_THROW(EXADG_ERR_HIP, hipGetErrorString(e));
  double sums[2];
  HIP_CHECK(hipMemcpyAsync(sums, err_scratch_.data() + 2 * 1024, sizeof(sums), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  return sums[1] > 0 ? std::sqrt(sums[0] / sums[1]) : std::sqrt(sums[0]);
}

std::unique_ptr<LaplaceOperatorBase> create_laplace_operator(exadg_mesh_desc const & mesh,
                                                             exadg_bc_desc const & bc, int degree,
                                                             int precision, double ip_factor)
{
  auto                grid = std::make_shared<StructuredGrid const>(mesh, bc);
  LaplaceOperatorData data;
  data.degree    = degree;
  data.IP_factor = ip_factor;
  if(precision == EXADG_FP64)
    return std::make_unique<LaplaceOperator<double>>(grid, data);
  if(precision == EXADG_FP32)
    return std::make_unique<LaplaceOperator<float>>(grid, data);
  EXADG_THROW(EXADG_ERR_ARGUMENT, "precision must be EXADG_FP64 or EXADG_FP32");
}

template class LaplaceOperator<double>;
template class CellBlockInverse<double>;
template class CellBlockInverse<float>;
template class LaplaceOperator<float>;

}  // namespace Poisson
}  // namespace ExaDG
