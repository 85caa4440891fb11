// vk_common.h -- shared definitions of the gfx950 library.
//
// Device layout is the reference's host layout, unchanged, so upload and
// download are flat copies: a halo field (0:nx+1,0:ny+1,0:nz+1) is a
// column-major double array (x fastest, strides 1, nx+2, (nx+2)(ny+2));
// cellvalue (1:nx,1:ny,1:nz,1:4) has no halo; the Poisson spectrum
// PHIT(0:nx/2,0:ny+1,0:nz+1) is interleaved complex double.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace vk {

struct Geom {
  int nx, ny, nz;
  long long sy, sz;
  __host__ __device__ Geom(int x, int y, int z)
      : nx(x), ny(y), nz(z), sy(x + 2), sz((long long)(x + 2) * (y + 2)) {}
  __host__ __device__ long long at(int i, int j, int k) const {
    return i + sy * j + sz * k;
  }
  // cellvalue(i,j,k,n), 1-based i,j,k,n
  __host__ __device__ long long cv(int i, int j, int k, int n) const {
    return (long long)(i - 1) +
           (long long)nx * ((j - 1) + (long long)ny * ((k - 1) + (long long)nz * (n - 1)));
  }
  __host__ __device__ size_t cells() const { return (size_t)sz * (size_t)(nz + 2); }
  __host__ __device__ size_t interior() const { return (size_t)nx * ny * nz; }
  // spectrum PHIT(0:nx/2,0:ny+1,0:nz+1)
  __host__ __device__ long long nh() const { return nx / 2; }
  __host__ __device__ long long spec_at(int i, int j, int k) const {
    return i + (nh() + 1) * (j + (long long)(ny + 2) * k);
  }
  __host__ __device__ size_t spec_cells() const {
    return (size_t)(nh() + 1) * (size_t)(ny + 2) * (size_t)(nz + 2);
  }
  // 32-bit element offsets, for kernels that hold many addresses at once
  // (an unsigned offset lets a load use a scalar base and one address
  // register); valid when fits32()
  __host__ __device__ bool fits32() const {
    const size_t lim = (size_t)1 << 32;
    return cells() < lim && 4 * interior() < lim && 2 * spec_cells() < lim;
  }
  __host__ __device__ unsigned at32(int i, int j, int k) const {
    return (unsigned)i + (unsigned)sy * (unsigned)j + (unsigned)sz * (unsigned)k;
  }
  __host__ __device__ unsigned cv32(int i, int j, int k, int n) const {
    return (unsigned)(i - 1) +
           (unsigned)nx * ((unsigned)(j - 1) + (unsigned)ny * ((unsigned)(k - 1) +
                                                               (unsigned)nz * (unsigned)(n - 1)));
  }
  __host__ __device__ unsigned spec32(int i, int j, int k) const {
    return (unsigned)i + (unsigned)(nh() + 1) * ((unsigned)j + (unsigned)(ny + 2) * (unsigned)k);
  }
};

struct Metrics {
  const double *dx, *dy, *dz, *dx12, *dy12, *dz12;
};

// Reciprocals 1/dx, 1/dy, ... of the metric arrays, for the kernels of the
// fused step that multiply where the reference divides.
struct RMetrics {
  const double *rdx, *rdy, *rdz, *rdx12, *rdy12, *rdz12;
};

// XCD-aware block order.  The dispatcher deals workgroups to the eight
// XCDs round-robin in linear block order; this renumbers them so each XCD
// works on one contiguous range of tiles and the neighbouring planes a
// stencil re-reads stay in that XCD's L2.  Returns the logical block
// (x, y, z) for the current workgroup.
__device__ inline void xcd_block(int& bx, int& by, int& bz) {
  const unsigned nb = gridDim.x * gridDim.y * gridDim.z;
  const unsigned p = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  const unsigned X = 8;
  const unsigned per = nb / X, rem = nb % X;
  const unsigned x = p % X, r = p / X;
  const unsigned q = x * per + (x < rem ? x : rem) + r;
  bx = (int)(q % gridDim.x);
  by = (int)((q / gridDim.x) % gridDim.y);
  bz = (int)(q / (gridDim.x * gridDim.y));
}

// decode halo index q into (i,j,k); returns false past the end.
// segments: z faces (all i,j), y faces (k interior, all i), x faces
// (j,k interior), and -- when `vplane` -- the interior of the j = ny plane.
__device__ inline bool halo_decode(const Geom& g, long long q, bool vplane, int& i, int& j,
                                   int& k) {
  const long long NX2 = g.nx + 2, NY2 = g.ny + 2;
  long long nzf = 2 * NX2 * NY2;
  if (q < nzf) {
    k = (q < NX2 * NY2) ? 0 : g.nz + 1;
    long long r = q % (NX2 * NY2);
    i = (int)(r % NX2);
    j = (int)(r / NX2);
    return true;
  }
  q -= nzf;
  long long nyf = 2 * NX2 * g.nz;
  if (q < nyf) {
    j = (q < NX2 * g.nz) ? 0 : g.ny + 1;
    long long r = q % (NX2 * g.nz);
    i = (int)(r % NX2);
    k = 1 + (int)(r / NX2);
    return true;
  }
  q -= nyf;
  long long nxf = 2LL * g.ny * g.nz;
  if (q < nxf) {
    long long r = q >> 1;
    i = (q & 1) ? g.nx + 1 : 0;
    j = 1 + (int)(r % g.ny);
    k = 1 + (int)(r / g.ny);
    return true;
  }
  q -= nxf;
  if (vplane && q < (long long)g.nx * g.nz) {
    j = g.ny;
    i = 1 + (int)(q % g.nx);
    k = 1 + (int)(q / g.nx);
    return true;
  }
  return false;
}

__host__ __device__ inline long long halo_count(const Geom& g, bool vplane) {
  long long n = 2LL * (g.nx + 2) * (g.ny + 2) + 2LL * (g.nx + 2) * g.nz + 2LL * g.ny * g.nz;
  return vplane ? n + (long long)g.nx * g.nz : n;
}

// A launch over the interior in blocks of kFlatThreads consecutive cells of
// the halo array (linear index c = i + sy j + sz k, for grids that fit32()),
// each block's first cell a multiple of kFlatThreads, in the dispatcher's
// own block order.  Every 128-byte line of a field is then written by one
// block: with 2D tiles the ends of a tile's row segments (rows start 8
// bytes into a line) share lines with the tiles beside them, which run on
// other XCDs, so those lines are written back in pieces from two L2s.
// Lanes on halo cells take no point.
constexpr int kFlatThreads = 256;
struct FlatRange {
  unsigned c0;  // first cell of block 0
  unsigned nb;  // blocks
};
inline FlatRange flat_interior(const Geom& g) {
  const unsigned first = g.at32(1, 1, 1), last = g.at32(g.nx, g.ny, g.nz);
  const unsigned c0 = first / kFlatThreads * kFlatThreads;
  return FlatRange{c0, (last - c0) / kFlatThreads + 1};
}
// cell c of a flat launch as (i, j, k); false on a halo cell or past the
// array (the last block's lanes run past it; c is 64-bit, so that they do
// not wrap round to cells of the array's start)
__device__ inline bool flat_point(const Geom& g, unsigned long long c, int& i, int& j, int& k) {
  if (c >= (unsigned long long)g.cells()) return false;
  const unsigned c32 = (unsigned)c;  // below cells(), so 32 bits hold it
  const unsigned sz = (unsigned)g.sz, sy = (unsigned)g.sy;
  const unsigned kk = c32 / sz, r = c32 - kk * sz;
  const unsigned jj = r / sy, ii = r - jj * sy;
  i = (int)ii;
  j = (int)jj;
  k = (int)kk;
  return ii - 1u < (unsigned)g.nx && jj - 1u < (unsigned)g.ny && kk - 1u < (unsigned)g.nz;
}
__device__ inline unsigned long long flat_cell(const FlatRange& fr) {
  return (unsigned long long)fr.c0 + (unsigned long long)blockIdx.x * kFlatThreads + threadIdx.x;
}

// A launch over the interior points (i,j,k) in 1..n, 64 x 4 threads per
// block along (i, j), one z plane per grid row.
constexpr int kBX = 64, kBY = 4;
inline dim3 interior_grid(const Geom& g) {
  return dim3((g.nx + kBX - 1) / kBX, (g.ny + kBY - 1) / kBY, g.nz);
}
inline dim3 interior_block() { return dim3(kBX, kBY, 1); }

}  // namespace vk
