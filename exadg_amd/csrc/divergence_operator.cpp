#include "divergence_operator.h"

#include <cmath>
#include <string>

namespace ExaDG
{
DivergenceOperator::DivergenceOperator(std::shared_ptr<StructuredGrid const> grid, int degree_p,
                                       int degree_u)
  : grid_(std::move(grid)), degree_p_(degree_p), degree_u_(degree_u)
{
  if(degree_u < 1 || degree_u > 8 || degree_p < 1 || degree_p > 8)
    EXADG_THROW(EXADG_ERR_ARGUMENT, "velocity and pressure degrees must be in [1,8]");
  ShapeInfo1D const su = make_shape_info(degree_u);
  ShapeInfo1D const sp = make_shape_info(degree_p);
  int const         nu = su.n, np = sp.n;
  nu3_                 = nu * nu * nu;
  np3_                 = np * np * np;
  // Su: velocity basis at the velocity Gauss points; Sp, Sdp: pressure basis and derivative there
  std::vector<double> tab(su.S);
  std::vector<double> val(np), der(np);
  std::vector<double> Sp((size_t)nu * np), Sdp((size_t)nu * np);
  for(int q = 0; q < nu; ++q)
  {
    lagrange_basis(sp.nodes, su.qp[q], val.data(), der.data());
    for(int i = 0; i < np; ++i)
    {
      Sp[q * np + i]  = val[i];
      Sdp[q * np + i] = der[i];
    }
  }
  tab.insert(tab.end(), Sp.begin(), Sp.end());
  tab.insert(tab.end(), Sdp.begin(), Sdp.end());
  tables_.upload(tab);
  // point data at the velocity quadrature (positions, JxW J^{-1}; face JxW n)
  DivergenceGeometry const g = compute_divergence_geometry(*grid_, su);
  cell_geo_.upload(g.cell);
  face_geo_.upload(g.face);
}

void DivergenceOperator::apply(double * dst, double const * src, hipStream_t s) const
{
  if(dst == nullptr || src == nullptr)
    EXADG_THROW(EXADG_ERR_ARGUMENT, "null vector");
  if((void const *)dst == (void const *)src)
    EXADG_THROW(EXADG_ERR_ARGUMENT, "dst and src must not alias");
  DivergenceArgs a;
  a.dst      = dst;
  a.src      = src;
  a.cell_geo = cell_geo_.data();
  a.face_geo = face_geo_.data();
  a.tables   = tables_.data();
  a.topo     = grid_->topology();
  a.n_u      = degree_u_ + 1;
  a.n_p      = degree_p_ + 1;
  a.n_cells  = grid_->n_cells();
  hipError_t const e = launch_divergence_apply(a, s);
  if(e != hipSuccess)
    EXADG_THROW(EXADG_ERR_HIP, std::string("divergence kernel launch: ") + hipGetErrorString(e));
}

void DivergenceOperator::ensure_nodes() const
{
  if(nodes_.size() == 0)
  {
    // positions of the Gauss-Lobatto nodes of every cell, component-major
    ShapeInfo1D const   su = make_shape_info(degree_u_);
    int64_t const       nc = grid_->n_cells(), np = nc * nu3_;
    std::vector<double> h((size_t)3 * np);
    parallel_range(nc, [&](int64_t b, int64_t e) {
      std::vector<double> xx(3 * nu3_), J(9 * nu3_);
      for(int64_t c = b; c < e; ++c)
      {
        grid_->evaluate(c, su.nodes, su.nodes, su.nodes, xx.data(), J.data());
        for(int d = 0; d < 3; ++d)
          for(int i = 0; i < nu3_; ++i)
            h[(size_t)d * np + c * nu3_ + i] = xx[d * nu3_ + i];
      }
    });
    nodes_.upload(h);
  }
}

void DivergenceOperator::velocity_points(double * x, hipStream_t s) const
{
  if(x == nullptr)
    EXADG_THROW(EXADG_ERR_ARGUMENT, "null vector");
  ensure_nodes();
  HIP_CHECK(hipMemcpyAsync(x, nodes_.data(), nodes_.size() * sizeof(double), hipMemcpyDeviceToDevice, s));
}

void DivergenceOperator::interpolate_fourier(double const * modes_host, int n_modes, double t,
                                             double * u, hipStream_t s) const
{
  if(u == nullptr || modes_host == nullptr || n_modes < 1)
    EXADG_THROW(EXADG_ERR_ARGUMENT, "invalid Fourier field");
  ensure_nodes();
  // c = A exp(i (phi + omega t)), as the divergence right-hand side of the Laplace operator
  std::vector<double> mh((size_t)8 * n_modes);
  int                 K = 0;
  for(int m = 0; m < n_modes; ++m)
  {
    double const * q = modes_host + 10 * m;
    for(int d = 0; d < 3; ++d)
    {
      if(q[d] != std::round(q[d]) || std::abs(q[d]) > 8)
        EXADG_THROW(EXADG_ERR_ARGUMENT, "wave vectors must be integer with |n_d| <= 8");
      K                 = std::max(K, (int)std::abs(q[d]));
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
  hipError_t const e = launch_fourier_interpolate(u, nodes_.data(), modes_.data(), n_modes, K, nu3_,
                                                  grid_->n_cells(), s);
  if(e != hipSuccess)
    EXADG_THROW(EXADG_ERR_HIP, std::string("interpolation kernel launch: ") + hipGetErrorString(e));
  HIP_CHECK(hipStreamSynchronize(s));  // the mode staging buffer is reused by the next call
}
}  // namespace ExaDG
