// tile.h -- 2D tiling of cell grids for the solver kernels.
//
// A workgroup of 256 threads covers a 64 x 4 tile (one wave per row
// segment: 64 consecutive cells, coalesced).  The hardware hands workgroup
// b to XCD b % 8; TileOf() renumbers workgroups so each XCD walks a
// contiguous band of rows, keeping the vertical stencil neighbours of a tile
// in the same L2.
#pragma once

#include <hip/hip_runtime.h>

namespace hg {

constexpr int kTX = 64;
constexpr int kTY = 4;
constexpr int kTileThreads = kTX * kTY;
constexpr int kXcds = 8;

struct TileGrid {
  int ntx, nty, per;  // tiles in x, y; tiles per XCD
  int blocks() const { return per * kXcds; }
};

inline TileGrid MakeTileGrid(int nx, int ny, int ty = kTY) {
  TileGrid t;
  t.ntx = (nx + kTX - 1) / kTX;
  t.nty = (ny + ty - 1) / ty;
  int nt = t.ntx * t.nty;
  t.per = (nt + kXcds - 1) / kXcds;
  return t;
}

// Cell of this thread; returns false for threads outside the grid (they
// still take part in block reductions with zero contributions).
__device__ __forceinline__ bool TileCell(const TileGrid& t, int nx, int ny, int& i, int& j) {
  const int b = blockIdx.x;
  const int tile = (b % kXcds) * t.per + b / kXcds;
  const int tx = tile % t.ntx, ty = tile / t.ntx;
  i = tx * kTX + (threadIdx.x & (kTX - 1));
  j = ty * kTY + (threadIdx.x >> 6);
  return tile < t.ntx * t.nty && i < nx && j < ny;
}

}  // namespace hg
