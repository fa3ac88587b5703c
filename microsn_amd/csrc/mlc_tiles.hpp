// mlc_tiles.hpp -- memory order of the node maps (convolved (mu, dmu/drad)
// pairs): 4 x 4 pixel tiles, tiles row-major over the map, pixels row-major
// inside a tile.  A tile of FP32 pairs is one 128-byte cache line, so the
// 2 x 2 footprint of a bilinear tap touches one line in most cases whatever
// the direction of the trajectory.  Nx and Ny must be multiples of 4.
#pragma once

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define MLC_THD __host__ __device__ __forceinline__
#else
#define MLC_THD inline
#endif

namespace mlc {

MLC_THD long long tile_offset(int x, int y, int Nx) {
  return ((long long)(y >> 2) * (Nx >> 2) + (x >> 2)) * 16 + ((y & 3) << 2) + (x & 3);
}

}  // namespace mlc
