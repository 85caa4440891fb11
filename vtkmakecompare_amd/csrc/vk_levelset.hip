// vk_levelset.hip -- set_levelset (set_levelset.f:1-199): 2001 sweeps of
// Godunov-upwind reinitialisation phi <- phi - dt sign(phi0)(|grad phi|-1)
// with zero-gradient halos, then the smoothed Heaviside cellvalue(...,1).
#include <cmath>
#include <string>
#include <vector>

#include "vk_kernels.h"
#include "vk_runtime.h"

namespace vk {

constexpr int kMaxShapes = 1 << 16;

// set_levelset.f:32-96 on the interior, reading `a`, writing `b`
__global__ void k_reinit(Geom g, const double* __restrict__ a, double* __restrict__ b,
                         const double* __restrict__ phi0, const double* dx, const double* dx12,
                         const double* dy12, const double* dz12) {
  int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  int k = 1 + blockIdx.z;
  if (i > g.nx || j > g.ny) return;
  const double dtl = 1.0e-3, epsilon = 1.0 - 1.0;  // PARAMETER(epsilon=1.0d0-1) is 0
  long long o = g.at(i, j, k);
  double c = a[o];
  double fa = (c - a[o - 1]) / dx12[i - 1];
  double fb = (a[o + 1] - c) / dx[i];
  double fc = (c - a[o - g.sy]) / dy12[j - 1];
  double fd = (a[o + g.sy] - c) / dy12[j];
  double fe = (c - a[o - g.sz]) / dz12[k - 1];
  double ff = (a[o + g.sz] - c) / dz12[k];
  double ap = fa >= 0.0 ? fa : 0.0, am = fa >= 0.0 ? 0.0 : fa;
  double bp = fb >= 0.0 ? fb : 0.0, bm = fb >= 0.0 ? 0.0 : fb;
  double cp = fc >= 0.0 ? fc : 0.0, cm = fc >= 0.0 ? 0.0 : fc;
  double dp = fd >= 0.0 ? fd : 0.0, dm = fd >= 0.0 ? 0.0 : fd;
  double ep = fe >= 0.0 ? fe : 0.0, em = fe >= 0.0 ? 0.0 : fe;
  double fp = ff >= 0.0 ? ff : 0.0, fm = ff >= 0.0 ? 0.0 : ff;
  double p0 = phi0[o];
  double sss = p0 / sqrt(p0 * p0 + epsilon * epsilon);
  double ggg;
  if (p0 > 0.0)
    ggg = sqrt(fmax(ap * ap, bm * bm) + fmax(cp * cp, dm * dm) + fmax(ep * ep, fm * fm)) - 1.0;
  else if (p0 < 0.0)
    ggg = sqrt(fmax(am * am, bp * bp) + fmax(cm * cm, dp * dp) + fmax(em * em, fp * fp)) - 1.0;
  else
    ggg = 0.0;
  b[o] = c - dtl * sss * ggg;
}

// zero-gradient halo in all three directions (set_levelset.f:107-124);
// every copy reads one interior cell, so one pass over the halo cells
// gives the same values
__global__ void k_halo_copy_all(Geom g, double* f) {
  long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  int i, j, k;
  if (!halo_decode(g, q, false, i, j, k)) return;
  int im = i == 0 ? 1 : (i == g.nx + 1 ? g.nx : i);
  int jm = j == 0 ? 1 : (j == g.ny + 1 ? g.ny : j);
  int km = k == 0 ? 1 : (k == g.nz + 1 ? g.nz : k);
  f[g.at(i, j, k)] = f[g.at(im, jm, km)];
}

// set_levelset.f:157-170, pi = 3.1415926535
__global__ void k_heaviside(Geom g, const double* phi, const double* dx, double* cv) {
  int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  int k = 1 + blockIdx.z;
  if (i > g.nx || j > g.ny) return;
  const double pi = 3.1415926535;
  double ph = phi[g.at(i, j, k)];
  double h;
  if (ph >= dx[i])
    h = 1.0;
  else if (ph <= -1.0 * dx[i])
    h = 0.0;
  else
    h = (sin(pi * ph / 2.0 / dx[i]) + 1.0) / 2.0;
  cv[g.cv(i, j, k, 1)] = h;
}

// One shape of the generator, as the kernel reads it: centre, reciprocal
// semi-axes, the rotation R (body axes in the columns, so body coordinates
// are R^T d) and the smallest semi-axis.
struct ShapeRec {
  double c[3];
  double ra[3];
  double r[9];  // row-major
  double amin;
};
static_assert(sizeof(ShapeRec) == 16 * sizeof(double), "ShapeRec is 16 doubles");

// phi0 = max over shapes of amin (1 - |R^T d / (a, b, c)|), positive inside
// (the solid), d the minimum-image offset from the centre in the periodic
// x and z; -1 with no shape.  Halo cells are set to 0, as set_levelset
// expects of a phi0 read by readgrid3d into a zeroed array.
__global__ void k_shapes(Geom g, const double* __restrict__ x, const double* __restrict__ y,
                         const double* __restrict__ z, double xl, double zl, int n,
                         const ShapeRec* __restrict__ sh, double* __restrict__ phi0) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  int k = blockIdx.z;
  if (i > g.nx + 1 || j > g.ny + 1) return;
  const long long o = g.at(i, j, k);
  if (i == 0 || j == 0 || k == 0 || i == g.nx + 1 || j == g.ny + 1 || k == g.nz + 1) {
    phi0[o] = 0.0;
    return;
  }
  const double px = x[i], py = y[j], pz = z[k];
  double best = -1.0;
  for (int q = 0; q < n; q++) {
    const ShapeRec& s = sh[q];
    double dx = px - s.c[0], dy = py - s.c[1], dz = pz - s.c[2];
    if (dx > 0.5 * xl) dx -= xl;
    if (dx < -0.5 * xl) dx += xl;
    if (dz > 0.5 * zl) dz -= zl;
    if (dz < -0.5 * zl) dz += zl;
    const double b0 = (s.r[0] * dx + s.r[3] * dy + s.r[6] * dz) * s.ra[0];
    const double b1 = (s.r[1] * dx + s.r[4] * dy + s.r[7] * dz) * s.ra[1];
    const double b2 = (s.r[2] * dx + s.r[5] * dy + s.r[8] * dz) * s.ra[2];
    const double v = s.amin * (1.0 - sqrt(b0 * b0 + b1 * b1 + b2 * b2));
    best = q == 0 ? v : fmax(best, v);
  }
  phi0[o] = best;
}

}  // namespace vk

using namespace vk;

extern "C" int vk_levelset_shapes(const int* nx, const int* ny, const int* nz, const double* x,
                                  const double* y, const double* z, const double* xl,
                                  const double* zl, const int* nshapes, const double* shapes,
                                  double* phi0) {
  VK_REQUIRE_INIT();
  Geom g(*nx, *ny, *nz);
  if (g.nx < 1 || g.ny < 1 || g.nz < 1) return fail(VK_ERR_BAD_SIZE, "levelset_shapes: n >= 1");
  const int n = *nshapes;
  if (n < 0 || n > kMaxShapes)
    return fail(VK_ERR_ARGUMENT, "levelset_shapes: 0 <= nshapes <= " +
                                     std::to_string(kMaxShapes));
  if (n > 0 && !shapes) return fail(VK_ERR_ARGUMENT, "levelset_shapes: shapes is null");
  if (!(*xl > 0.0) || !(*zl > 0.0)) return fail(VK_ERR_ARGUMENT, "levelset_shapes: xl, zl > 0");
  VK_RESOLVE(dphi, double, phi0, g.cells() * sizeof(double));
  VK_RESOLVE(dx, const double, x, (g.nx + 2) * sizeof(double));
  VK_RESOLVE(dy, const double, y, (g.ny + 2) * sizeof(double));
  VK_RESOLVE(dz, const double, z, (g.nz + 2) * sizeof(double));
  // host records: centre, semi-axes, unit quaternion (w, x, y, z)
  std::vector<ShapeRec> rec((size_t)(n > 0 ? n : 1));
  for (int q = 0; q < n; q++) {
    const double* p = shapes + 10 * (size_t)q;
    const double a0 = p[3], a1 = p[4], a2 = p[5];
    if (!(a0 > 0.0) || !(a1 > 0.0) || !(a2 > 0.0))
      return fail(VK_ERR_ARGUMENT, "levelset_shapes: semi-axes must be positive");
    double w = p[6], qx = p[7], qy = p[8], qz = p[9];
    const double nq = std::sqrt(w * w + qx * qx + qy * qy + qz * qz);
    if (!(nq > 0.0)) return fail(VK_ERR_ARGUMENT, "levelset_shapes: zero quaternion");
    w /= nq; qx /= nq; qy /= nq; qz /= nq;
    ShapeRec& r = rec[q];
    r.c[0] = p[0]; r.c[1] = p[1]; r.c[2] = p[2];
    r.ra[0] = 1.0 / a0; r.ra[1] = 1.0 / a1; r.ra[2] = 1.0 / a2;
    r.amin = std::fmin(a0, std::fmin(a1, a2));
    // rotation matrix of the unit quaternion, row-major
    r.r[0] = 1.0 - 2.0 * (qy * qy + qz * qz);
    r.r[1] = 2.0 * (qx * qy - qz * w);
    r.r[2] = 2.0 * (qx * qz + qy * w);
    r.r[3] = 2.0 * (qx * qy + qz * w);
    r.r[4] = 1.0 - 2.0 * (qx * qx + qz * qz);
    r.r[5] = 2.0 * (qy * qz - qx * w);
    r.r[6] = 2.0 * (qx * qz - qy * w);
    r.r[7] = 2.0 * (qy * qz + qx * w);
    r.r[8] = 1.0 - 2.0 * (qx * qx + qy * qy);
  }
  ShapeRec* dsh = static_cast<ShapeRec*>(workspace(kWorkMisc, sizeof(ShapeRec) * rec.size()));
  if (!dsh) return VK_ERR_HIP;
  hipStream_t s = rt().stream;
  int rc = hip_check(hipMemcpyAsync(dsh, rec.data(), sizeof(ShapeRec) * rec.size(),
                                    hipMemcpyHostToDevice, s),
                     "levelset_shapes upload");
  if (rc) return rc;
  // the host records must outlive the copy
  rc = hip_check(hipStreamSynchronize(s), "levelset_shapes upload");
  if (rc) return rc;
  {
    ProfScope ps("levelset_shapes", s);
    dim3 blk(64, 4);
    dim3 grd((unsigned)((g.nx + 2 + 63) / 64), (unsigned)((g.ny + 2 + 3) / 4),
             (unsigned)(g.nz + 2));
    k_shapes<<<grd, blk, 0, s>>>(g, dx, dy, dz, *xl, *zl, n, dsh, dphi);
  }
  VK_LAUNCH_CHECK("levelset_shapes");
  return VK_OK;
}

extern "C" int vk_set_levelset(const int* nx, const int* ny, const int* nz, const double* dx,
                               const double* dx12, const double* dy12, const double* dz12,
                               const double* phi0, double* phi, const double* x,
                               const double* y, const double* z, double* cellvalue) {
  VK_REQUIRE_INIT();
  (void)x; (void)y; (void)z;
  Geom g(*nx, *ny, *nz);
  if (g.nx < 2 || g.ny < 2 || g.nz < 2) return fail(VK_ERR_BAD_SIZE, "set_levelset: n >= 2");
  size_t fb = g.cells() * sizeof(double);
  VK_RESOLVE(dp0, const double, phi0, fb);
  VK_RESOLVE(dphi, double, phi, fb);
  VK_RESOLVE(dcv, double, cellvalue, g.interior() * 4 * sizeof(double));
  VK_RESOLVE(ddx, const double, dx, (g.nx + 2) * sizeof(double));
  VK_RESOLVE(ddx12, const double, dx12, (g.nx + 2) * sizeof(double));
  VK_RESOLVE(ddy12, const double, dy12, (g.ny + 2) * sizeof(double));
  VK_RESOLVE(ddz12, const double, dz12, (g.nz + 2) * sizeof(double));
  double* work = static_cast<double*>(workspace(kWorkPhi, fb));
  if (!work) return VK_ERR_HIP;
  hipStream_t s = rt().stream;
  int rc = hip_check(hipMemcpyAsync(dphi, dp0, fb, hipMemcpyDeviceToDevice, s), "levelset copy");
  if (rc) return rc;
  rc = hip_check(hipMemcpyAsync(work, dp0, fb, hipMemcpyDeviceToDevice, s), "levelset copy");
  if (rc) return rc;
  const int iterations = 2001;  // it = 0..nnn with nnn = 2000 (set_levelset.f:4,30)
  long long tot = halo_count(g, false);
  double* a = dphi;
  double* b = work;
  {
    ProfScope ps("levelset_reinit", s);  // the 2001 iterations, timed as one
    for (int it = 0; it < iterations; it++) {
      k_reinit<<<interior_grid(g), interior_block(), 0, s>>>(g, a, b, dp0, ddx, ddx12, ddy12,
                                                             ddz12);
      k_halo_copy_all<<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(g, b);
      double* tmp = a;
      a = b;
      b = tmp;
    }
  }
  if (a != dphi) {
    rc = hip_check(hipMemcpyAsync(dphi, a, fb, hipMemcpyDeviceToDevice, s), "levelset copy");
    if (rc) return rc;
  }
  k_heaviside<<<interior_grid(g), interior_block(), 0, s>>>(g, dphi, ddx, dcv);
  VK_LAUNCH_CHECK("set_levelset");
  return VK_OK;
}
