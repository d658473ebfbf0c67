// Structured-box topology shared by host and device code: cell coordinates, neighbours
// and the global face numbering used by the per-face data arrays.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace ExaDG
{
struct BoxTopology
{
  int     nc[3];        // cells per direction
  int     periodic[3];
  int     layers[3];    // face layers normal to d: nc[d] (periodic) or nc[d]+1
  int     bc[6];        // boundary type per side (EXADG_BC_*)
  int64_t face_base[3]; // first face index of direction d
  int64_t n_faces;
};

// tangential directions of a face normal to d, in increasing order
__host__ __device__ inline int tangent1(int d) { return d == 0 ? 1 : 0; }
__host__ __device__ inline int tangent2(int d) { return d == 2 ? 1 : 2; }

__host__ __device__ inline void cell_coords(BoxTopology const & t, int64_t c, int * cc)
{
  cc[0]             = (int)(c % t.nc[0]);
  int64_t const r   = c / t.nc[0];
  cc[1]             = (int)(r % t.nc[1]);
  cc[2]             = (int)(r / t.nc[1]);
}

__host__ __device__ inline int64_t cell_id(BoxTopology const & t, int const * cc)
{
  return cc[0] + (int64_t)t.nc[0] * (cc[1] + (int64_t)t.nc[1] * cc[2]);
}

// neighbour of the cell at cc across face (d, side); -1 on a non-periodic boundary
__host__ __device__ inline int64_t neighbor_id(BoxTopology const & t, int const * cc, int d,
                                               int side)
{
  int n[3] = {cc[0], cc[1], cc[2]};
  n[d] += side ? 1 : -1;
  if(n[d] < 0 || n[d] >= t.nc[d])
  {
    if(!t.periodic[d])
      return -1;
    n[d] = (n[d] + t.nc[d]) % t.nc[d];
  }
  return cell_id(t, n);
}

// global index of face (d, side) of the cell at cc
__host__ __device__ inline int64_t face_id(BoxTopology const & t, int const * cc, int d, int side)
{
  int fl = cc[d] + side;
  if(t.periodic[d] && fl == t.nc[d])
    fl = 0;
  int const a = tangent1(d), b = tangent2(d);
  return t.face_base[d] + fl + (int64_t)t.layers[d] * (cc[a] + (int64_t)t.nc[a] * cc[b]);
}

}  // namespace ExaDG
