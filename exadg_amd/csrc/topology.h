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
    n[d] += n[d] < 0 ? t.nc[d] : -t.nc[d];  // one step across the periodic seam
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

// Cell visited at traversal position idx: tiles of `tile_rows` y-rows, each walked with x
// fastest, then y within the tile, then z, so that all six face neighbours of a cell are
// visited within about nx*tile_rows positions of it.
__host__ __device__ inline int64_t traversal_cell(BoxTopology const & t, int64_t idx, int tile_rows)
{
  int const     nx   = t.nc[0], ny = t.nc[1], nz = t.nc[2];
  int64_t const full = (int64_t)nx * nz * tile_rows;
  int64_t const tile = idx / full;
  int64_t const r    = idx - tile * full;
  int const     rows = ny - (int)tile * tile_rows < tile_rows ? ny - (int)tile * tile_rows : tile_rows;
  int const     x    = (int)(r % nx);
  int64_t const r2   = r / nx;
  int const     yi   = (int)(r2 % rows);
  int const     z    = (int)(r2 / rows);
  return x + (int64_t)nx * ((int64_t)tile * tile_rows + yi + (int64_t)ny * z);
}

// Coordinates of the cell at traversal position pos0 + slot (the order of traversal_cell). The
// divisions act on pos0 only, which the kernels keep uniform over a workgroup (scalar
// arithmetic); per lane there is at most a short carry along x.
__host__ __device__ inline void traversal_coords(BoxTopology const & t, int tile_rows, int pos0,
                                                 int slot, int (&cc)[3])
{
  int const nx = t.nc[0], ny = t.nc[1], nz = t.nc[2];
  int const full = nx * nz * tile_rows;
  int       tile = pos0 / full;
  int const r    = pos0 - tile * full;
  int       rows = ny - tile * tile_rows < tile_rows ? ny - tile * tile_rows : tile_rows;
  int       x    = r % nx;
  int const r2   = r / nx;
  int       yi   = r2 % rows;
  int       z    = r2 / rows;
  x += slot;
  while(x >= nx)
  {
    x -= nx;
    if(++yi == rows)
    {
      yi = 0;
      if(++z == nz)
      {
        z = 0;
        ++tile;
        rows = ny - tile * tile_rows < tile_rows ? ny - tile * tile_rows : tile_rows;
      }
    }
  }
  cc[0] = x;
  cc[1] = tile * tile_rows + yi;
  cc[2] = z;
}

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X (block b runs on XCD
// b % 8). Renumber them so that each XCD receives one contiguous range of logical blocks and
// therefore one contiguous part of the traversal (its own L2 serves the neighbour reads).
__host__ __device__ inline int64_t xcd_logical_block(int64_t b, int64_t n_blocks)
{
  int64_t const q = n_blocks / 8, r = n_blocks % 8;
  int64_t const x = b % 8, i = b / 8;
  return x * q + (x < r ? x : r) + i;
}

}  // namespace ExaDG
