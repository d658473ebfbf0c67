// Weak DG divergence of a discrete velocity (the right-hand side of the pressure step) and the
// nodal interpolation of a Fourier velocity field, FP64 (divergence_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "topology.h"

namespace ExaDG
{
struct DivergenceArgs
{
  double *       dst;        // pressure DoFs, n_p^3 per cell
  double const * src;        // velocity DoFs, 3 component blocks of n_u^3 per cell
  double const * cell_geo;   // per cell 12 x n_u^3: x (3), C = JxW J^{-1} (9), at the Gauss points
  double const * face_geo;   // per face 6 x n_u^2: x (3), JxW n (3), n the unit normal in +xi_d
  double const * tables;     // Su (n_u x n_u) | Sp (n_u x n_p) | Sdp (n_u x n_p)
  BoxTopology    topo;
  int            n_u, n_p;
  int64_t        n_cells;
};

// dst = B src: -(u, grad q)_K + sum_F ({u}.n_K, q)_F with the central flux; boundary faces:
// pressure Neumann sides are walls at rest (u+ = -u-: no flux), pressure Dirichlet sides
// velocity Neumann (u+ = u-: flux u-.n)
hipError_t launch_divergence_apply(DivergenceArgs const & a, hipStream_t s);

// u[c][k][node] = sum_m Re(d_m,k c_m exp(i 2 pi n_m.x_node)) at the given node positions
// (3 x n_points, component-major); mode_data: 8 doubles per mode n(3), d(3), c_re, c_im,
// |n_d| <= K <= 16. Nodes are numbered cell by cell (n3 per cell), the output in the velocity
// layout (per cell three component blocks).
hipError_t launch_fourier_interpolate(double * u, double const * x_nodes, double const * mode_data,
                                      int n_modes, int K, int n3, int64_t n_cells, hipStream_t s);
}  // namespace ExaDG
