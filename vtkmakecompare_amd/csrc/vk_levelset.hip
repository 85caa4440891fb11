// vk_levelset.hip -- set_levelset (set_levelset.f:1-199): 2001 sweeps of
// Godunov-upwind reinitialisation phi <- phi - dt sign(phi0)(|grad phi|-1)
// with zero-gradient halos, then the smoothed Heaviside cellvalue(...,1).
#include "vk_kernels.h"
#include "vk_runtime.h"

namespace vk {

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

}  // namespace vk

using namespace vk;

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
