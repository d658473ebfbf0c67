// Launch interface of the error-norm kernel (see error_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "topology.h"

namespace ExaDG
{
// Per-cell partial sums of the L2 error against u* = prod_d sin(omega x_d) and of |u*|^2 with
// QGauss(ne): partial[2c] = sum_q JxW (u_h - u*)^2, partial[2c+1] = sum_q JxW u*^2.
// Se: ne x n values of the nodal basis at the error points; w: ne weights then ne points on
// [0,1]. Affine meshes: x and JxW from lower and h; curved meshes: point data (4 x ne^3 per
// cell: x, y, z, JxW) in geo.
hipError_t launch_error_sine(double * partial, double const * u, double omega, int n, int ne,
                             double const * Se, double const * w, BoxTopology const & topo,
                             double const * lower, double const * h, double const * geo,
                             int64_t n_cells, hipStream_t s);
// out[0..1] = column sums of partial (n_rows x 2), deterministic two-stage reduction
hipError_t launch_sum_pairs(double const * partial, int64_t n_rows, double * scratch,
                            double * out, hipStream_t s);
}  // namespace ExaDG
