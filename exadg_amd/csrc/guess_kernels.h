// Launch interface of the learned-initial-guess kernels (see guess_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "topology.h"

namespace ExaDG
{
// up to four previous solutions, newest first (passed by value to the kernels)
template<typename T>
struct History
{
  T const * p[4] = {nullptr, nullptr, nullptr, nullptr};
};

// reciprocal feature scales: backward differences of orders 0..3 of the history, rhs, target
struct FeatureScales
{
  float inv_d[4] = {1.f, 1.f, 1.f, 1.f};
  float inv_b    = 1.f;
  float inv_t    = 1.f;
};

// C(i,j) = alpha sum_k A(i,k) B(k,j) (+ C) (+ bias_j) (ReLU), A(i,k) = A[i a_i + k a_k],
// B(k,j) = B[k b_k + j b_j]; with a workspace of >= s*M*N floats the K range may be split
// over s blocks when M x N alone has too few tiles
hipError_t launch_sgemm(int M, int N, int K, float const * A, int64_t a_i, int64_t a_k,
                        float const * B, int64_t b_k, int64_t b_j, float * C, int64_t ldc,
                        float alpha, bool accumulate, float const * bias, bool relu,
                        float * workspace, int64_t workspace_floats, hipStream_t s);
template<typename In>
hipError_t launch_features(float * X, float * T, History<In> const & hist, In const * rhs,
                           In const * target, int history, int n_cell_dofs, int64_t cell0, int B,
                           FeatureScales const & sc, double const * beta, hipStream_t s);
// acc[g] += sum of squares of feature group g (history differences 0..H-1, rhs, target) over
// B samples; partial: >= 1024 * 6 doubles
hipError_t launch_feature_moments(float const * X, float const * T, int history, int n_cell_dofs,
                                  int B, double * partial, double * acc, hipStream_t s);
// dY = 2 scale (Y - T); loss_sum (+)= sum (Y - T)^2
hipError_t launch_mse_backward(float const * Y, float const * T, float * dY, int64_t n, float scale,
                               double * partial, double * loss_sum, bool accumulate, hipStream_t s);
hipError_t launch_relu_backward(float * dH, float const * H, int64_t n, hipStream_t s);
hipError_t launch_column_sum(float const * dY, int M, int N, float * db, hipStream_t s);
hipError_t launch_adam(float * p, float const * g, float * m, float * v, int64_t n, float lr,
                       float b1, float b2, float eps, int step, hipStream_t s);
// x0 = sum_h beta_h hist_h (+ st * Y) (- target); counts the non-finite entries of Y into n_bad
hipError_t launch_combine(double * x0, History<double> const & hist, int history,
                          double const * beta, float const * Y, double st, double const * target,
                          int64_t n, unsigned int * n_bad, hipStream_t s);
// dst = (float)(coef * src)
hipError_t launch_scale_to_float(float * dst, double const * src, double coef, int64_t n,
                                 hipStream_t s);
hipError_t launch_fill_u32(unsigned int * p, unsigned int v, int n, hipStream_t s);
// weak DG divergence (central flux, no-slip walls) of u(x) = Re sum_m d_m c_m exp(i 2 pi n_m.x);
// mode_data (device): 8 doubles per mode n_x, n_y, n_z, d[3], c_re, c_im with |n_d| <= K <= 16;
// S, Sd: n x n value/derivative tables at the Gauss points; w: n weights then n points
// the same on a curved mesh from compute_divergence_geometry's point data (K <= 16; LDS
// 8 (5 n^3 + 768 (K+1)) bytes)
hipError_t launch_divergence_rhs_curved(double * b, double const * mode_data, int n_modes, int K,
                                        BoxTopology const & topo, int n, double const * S,
                                        double const * Sd, double const * w,
                                        double const * cell_geo, double const * face_geo,
                                        int64_t n_cells, hipStream_t s);
hipError_t launch_divergence_rhs(double * b, double const * mode_data, int n_modes, int K,
                                 BoxTopology const & topo, double const * lower, double const * h,
                                 int n, double const * S, double const * Sd, double const * w,
                                 int64_t n_cells, hipStream_t s);
}  // namespace ExaDG
