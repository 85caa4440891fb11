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
};

struct Metrics {
  const double *dx, *dy, *dz, *dx12, *dy12, *dz12;
};

// A launch over the interior points (i,j,k) in 1..n, 64 x 4 threads per
// block along (i, j), one z plane per grid row.
constexpr int kBX = 64, kBY = 4;
inline dim3 interior_grid(const Geom& g) {
  return dim3((g.nx + kBX - 1) / kBX, (g.ny + kBY - 1) / kBY, g.nz);
}
inline dim3 interior_block() { return dim3(kBX, kBY, 1); }

}  // namespace vk
