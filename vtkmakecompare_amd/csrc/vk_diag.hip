// vk_diag.hip -- the reference's run-time diagnostics as GPU reductions:
// CHECK_COURANT (nse_dominik.f:2519-2601), MONITOR (:2797-2872) and
// CHECK_CONVERGENCE (:2875-2892).
//
// Each is a deterministic two-pass reduction: a fixed grid of blocks walks
// the interior in linear order, every block writes one partial, and one
// block combines the partials in block order.  Maxima (Courant, the
// largest |T - T_before| and its position) are exact; the sums are formed
// in that tree order instead of the reference's serial loop order, so they
// agree with it to rounding (a few ulp of the sum).
#include <cmath>
#include <cstdio>
#include <string>

#include "vk_kernels.h"
#include "vk_runtime.h"

namespace vk {

constexpr int kDiagBlocks = 1024;
constexpr int kDiagThreads = 256;

// block-wide combine of one value per thread (fixed order: the lanes of a
// wave by xor steps, then the waves in order)
template <typename T, typename Op>
__device__ inline T block_combine(T v, Op op, T* sh) {
  for (int off = 32; off > 0; off >>= 1) v = op(v, __shfl_xor(v, off));
  const int w = threadIdx.x / 64, nw = blockDim.x / 64;
  if (threadIdx.x % 64 == 0) sh[w] = v;
  __syncthreads();
  T r = sh[0];
  for (int q = 1; q < nw; q++) r = op(r, sh[q]);
  __syncthreads();
  return r;
}

struct Sum {
  __device__ double operator()(double a, double b) const { return a + b; }
};
// the reference's "if (m < c) m = c" from m = 0: a NaN never enters
struct Max0 {
  __device__ double operator()(double a, double b) const { return b > a ? b : a; }
};

__device__ inline void interior_ijk(const Geom& g, long long q, int& i, int& j, int& k) {
  i = 1 + (int)(q % g.nx);
  const long long r = q / g.nx;
  j = 1 + (int)(r % g.ny);
  k = 1 + (int)(r / g.ny);
}

// CHECK_COURANT: per block the largest u dt/dx, v dt/dy, w dt/dz (each
// ratio formed as (x dt) / d), from 0
__global__ __launch_bounds__(kDiagThreads) void k_courant(Geom g, const double* u,
                                                          const double* v, const double* w,
                                                          const double* dx, const double* dy,
                                                          const double* dz, double dt,
                                                          double* part) {
  __shared__ double sh[kDiagThreads / 64];
  const long long n = (long long)g.interior();
  double mx = 0.0, my = 0.0, mz = 0.0;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (long long)gridDim.x * blockDim.x) {
    int i, j, k;
    interior_ijk(g, q, i, j, k);
    const long long o = g.at(i, j, k);
    const double cx = u[o] * dt / dx[i], cy = v[o] * dt / dy[j], cz = w[o] * dt / dz[k];
    mx = cx > mx ? cx : mx;
    my = cy > my ? cy : my;
    mz = cz > mz ? cz : mz;
  }
  mx = block_combine(mx, Max0{}, sh);
  my = block_combine(my, Max0{}, sh);
  mz = block_combine(mz, Max0{}, sh);
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x] = mx;
    part[3 * blockIdx.x + 1] = my;
    part[3 * blockIdx.x + 2] = mz;
  }
}

__global__ __launch_bounds__(1024) void k_courant_final(const double* part, int nb,
                                                        double* out) {
  __shared__ double sh[16];
  double m[3] = {0.0, 0.0, 0.0};
  for (int b = threadIdx.x; b < nb; b += blockDim.x)
    for (int c = 0; c < 3; c++) m[c] = part[3 * b + c] > m[c] ? part[3 * b + c] : m[c];
  for (int c = 0; c < 3; c++) m[c] = block_combine(m[c], Max0{}, sh);
  if (threadIdx.x == 0) {
    // max(courant_max_x, courant_max_y, courant_max_z)
    double r = m[0];
    if (m[1] > r) r = m[1];
    if (m[2] > r) r = m[2];
    out[0] = r;
  }
}

// MONITOR, pass 1: per block the three velocity integrals and the largest
// |T - T_before| with its first linear position (k, j, i order)
// larger value, or the earlier position on a tie
__device__ inline void combine_pair(double& e, long long& q, double e2, long long q2) {
  if (e2 > e || (e2 == e && q2 < q)) {
    e = e2;
    q = q2;
  }
}

__global__ __launch_bounds__(kDiagThreads) void k_monitor(Geom g, const double* u,
                                                          const double* v, const double* w,
                                                          const double* t, const double* tb,
                                                          const double* dx, const double* dy,
                                                          const double* dz, double* part) {
  __shared__ double sh[kDiagThreads / 64];
  __shared__ double she[kDiagThreads / 64];
  __shared__ long long shq[kDiagThreads / 64];
  const long long n = (long long)g.interior();
  double su = 0.0, sv = 0.0, sw = 0.0, e = 0.0;
  long long qe = n;  // no position until a difference above 0 is seen
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (long long)gridDim.x * blockDim.x) {
    int i, j, k;
    interior_ijk(g, q, i, j, k);
    const long long o = g.at(i, j, k);
    su = su + fabs(u[o]) * dx[i];
    sv = sv + fabs(v[o]) * dy[j];
    sw = sw + fabs(w[o]) * dz[k];
    const double d = fabs(t[o] - tb[o]);
    if (d > e) {  // strict, as the reference: a NaN never enters
      e = d;
      qe = q;
    }
  }
  // T_all's share of this block: the plane k = nz + 1 (the k loop's index
  // after it ends, nse_dominik.f:2838-2843), in the same linear split
  const int kt = g.nz + 1;
  const long long np = (long long)g.nx * g.ny;
  double ta = 0.0;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < np;
       q += (long long)gridDim.x * blockDim.x) {
    const int i = 1 + (int)(q % g.nx), j = 1 + (int)(q / g.nx);
    ta = ta + t[g.at(i, j, kt)] * dx[i] * dy[j] * dz[kt];
  }
  su = block_combine(su, Sum{}, sh);
  sv = block_combine(sv, Sum{}, sh);
  sw = block_combine(sw, Sum{}, sh);
  ta = block_combine(ta, Sum{}, sh);
  for (int off = 32; off > 0; off >>= 1)
    combine_pair(e, qe, __shfl_xor(e, off), __shfl_xor(qe, off));
  if (threadIdx.x % 64 == 0) {
    she[threadIdx.x / 64] = e;
    shq[threadIdx.x / 64] = qe;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 1; q < kDiagThreads / 64; q++) combine_pair(e, qe, she[q], shq[q]);
    double* p = part + 6 * blockIdx.x;
    p[0] = su;
    p[1] = sv;
    p[2] = sw;
    p[3] = e;
    p[4] = (double)qe;  // exact below 2^53
    p[5] = ta;
  }
}

// MONITOR, pass 2: the partials in block order
__global__ __launch_bounds__(1024) void k_monitor_final(Geom g, const double* part, int nb,
                                                        double* out) {
  __shared__ double sh[16];
  __shared__ double she[16];
  __shared__ long long shq[16];
  const long long n = (long long)g.interior();
  double su = 0.0, sv = 0.0, sw = 0.0, e = 0.0, ta = 0.0;
  long long qe = n;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    const double* p = part + 6 * b;
    su = su + p[0];
    sv = sv + p[1];
    sw = sw + p[2];
    combine_pair(e, qe, p[3], (long long)p[4]);
    ta = ta + p[5];
  }
  su = block_combine(su, Sum{}, sh);
  sv = block_combine(sv, Sum{}, sh);
  sw = block_combine(sw, Sum{}, sh);
  ta = block_combine(ta, Sum{}, sh);
  for (int off = 32; off > 0; off >>= 1)
    combine_pair(e, qe, __shfl_xor(e, off), __shfl_xor(qe, off));
  if (threadIdx.x % 64 == 0) {
    she[threadIdx.x / 64] = e;
    shq[threadIdx.x / 64] = qe;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 1; q < (int)(blockDim.x / 64); q++) combine_pair(e, qe, she[q], shq[q]);
    out[0] = su;
    out[1] = sv;
    out[2] = sw;
    out[3] = ta;
    out[4] = e;
    out[5] = (double)qe;
  }
}

// CHECK_CONVERGENCE, pass 1: block b sums the planes k = 1 + b, 1 + b +
// gridDim.x, ... of the line i = nx (the threads over j)
__global__ __launch_bounds__(kDiagThreads) void k_convergence(Geom g, const double* u,
                                                              const double* y, double yl,
                                                              double re, double* part) {
  __shared__ double sh[kDiagThreads / 64];
  double s = 0.0;
  for (int k = 1 + blockIdx.x; k <= g.nz; k += gridDim.x)
    for (int j = 1 + threadIdx.x; j <= g.ny; j += blockDim.x) {
      const double a = -(re / 2.0 * y[j] * (y[j] - yl));
      s = s + fabs(u[g.at(g.nx, j, k)] - a);
    }
  s = block_combine(s, Sum{}, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// pass 2: the partials in block order
__global__ __launch_bounds__(1024) void k_sum_final(const double* part, int nb, double* out) {
  __shared__ double sh[16];
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) s = s + part[b];
  s = block_combine(s, Sum{}, sh);
  if (threadIdx.x == 0) out[0] = s;
}

// the device results, copied back and waited for
static int fetch(const double* dres, double* host, int n, hipStream_t s) {
  int rc = hip_check(hipMemcpyAsync(host, dres, n * sizeof(double), hipMemcpyDeviceToHost, s),
                     "diagnostics copy");
  return rc ? rc : hip_check(hipStreamSynchronize(s), "diagnostics sync");
}

static int diag_blocks(const Geom& g) {
  const long long nb = ((long long)g.interior() + kDiagThreads - 1) / kDiagThreads;
  return (int)(nb < kDiagBlocks ? nb : kDiagBlocks);
}

// The scratch of vk_check_courant_noalloc: the result (slot 0, then 7
// spare slots) and three partial maxima per block of the Courant pass.
// The pass's grid is capped at kDiagBlocks blocks, so 24.6 KB serves every
// grid.  Each device holds its own copy until the process ends.
constexpr int kCourantScratch = 8 + 3 * kDiagBlocks;
static __device__ __align__(256) double scratch[kCourantScratch];

// `scratch` is a host-side stand-in in host code, not a device address:
// its address on each device, looked up once per device (with that device
// current) and used for every kernel argument and copy
constexpr int kMaxDevices = 64;
static double* g_scratch[kMaxDevices] = {};

static double* courant_scratch() {
  int cur = -1;
  if (hip_check(hipGetDevice(&cur), "check_courant_noalloc: hipGetDevice")) return nullptr;
  if (cur != rt().device || cur < 0 || cur >= kMaxDevices) {
    fail(VK_ERR_ARGUMENT, "check_courant_noalloc: the current device is not the library's");
    return nullptr;
  }
  if (!g_scratch[cur]) {
    void* p = nullptr;
    if (hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)),
                  "check_courant_noalloc: hipGetSymbolAddress"))
      return nullptr;
    g_scratch[cur] = static_cast<double*>(p);
  }
  return g_scratch[cur];
}

}  // namespace vk

using namespace vk;

#define DIAG_GRID()                                                         \
  Geom g(*nx, *ny, *nz);                                                    \
  if (g.nx < 1 || g.ny < 1 || g.nz < 1)                                     \
    return fail(VK_ERR_BAD_SIZE, "diagnostics: nx, ny, nz >= 1");           \
  if ((double)g.interior() >= 9007199254740992.0)                           \
    return fail(VK_ERR_BAD_SIZE, "diagnostics: grid too large")
#define DIAG_FIELD(var, ptr) VK_RESOLVE(var, const double, ptr, g.cells() * sizeof(double))

extern "C" {

int vk_check_courant(const double* dt, const double* dx, const double* dy, const double* dz,
                     const int* nx, const int* ny, const int* nz, const double* u,
                     const double* v, const double* w, double* courant) {
  VK_REQUIRE_INIT();
  DIAG_GRID();
  DIAG_FIELD(du, u);
  DIAG_FIELD(dv, v);
  DIAG_FIELD(dw, w);
  VK_RESOLVE(ddx, const double, dx, (g.nx + 2) * sizeof(double));
  VK_RESOLVE(ddy, const double, dy, (g.ny + 2) * sizeof(double));
  VK_RESOLVE(ddz, const double, dz, (g.nz + 2) * sizeof(double));
  if (!courant) return fail(VK_ERR_ARGUMENT, "check_courant: courant is null");
  const int nb = diag_blocks(g);
  double* ws = static_cast<double*>(workspace(kWorkDiag, (size_t)(6 * kDiagBlocks + 8) *
                                                             sizeof(double)));
  if (!ws) return VK_ERR_HIP;
  hipStream_t s = rt().stream;
  {
    ProfScope ps("check_courant", s);
    k_courant<<<nb, kDiagThreads, 0, s>>>(g, du, dv, dw, ddx, ddy, ddz, *dt, ws + 8);
    k_courant_final<<<1, 1024, 0, s>>>(ws + 8, nb, ws);
  }
  VK_LAUNCH_CHECK("check_courant");
  return fetch(ws, courant, 1, s);
}

int vk_check_courant_noalloc(const double* dt, const double* dx, const double* dy,
                             const double* dz, const int* nx, const int* ny, const int* nz,
                             const double* u, const double* v, const double* w,
                             double* courant) {
  VK_REQUIRE_INIT();
  DIAG_GRID();
  DIAG_FIELD(du, u);
  DIAG_FIELD(dv, v);
  DIAG_FIELD(dw, w);
  VK_RESOLVE(ddx, const double, dx, (g.nx + 2) * sizeof(double));
  VK_RESOLVE(ddy, const double, dy, (g.ny + 2) * sizeof(double));
  VK_RESOLVE(ddz, const double, dz, (g.nz + 2) * sizeof(double));
  if (!courant) return fail(VK_ERR_ARGUMENT, "check_courant_noalloc: courant is null");
  const int nb = diag_blocks(g);
  if (8 + 3 * nb > kCourantScratch)
    return fail(VK_ERR_BAD_SIZE, "check_courant_noalloc: " + std::to_string(nb) +
                                     " partials exceed the scratch of " +
                                     std::to_string(kCourantScratch) + " doubles");
  double* ws = courant_scratch();
  if (!ws) return last_status();
  // every partial the final pass reads is written by the first pass, and
  // the result by the final pass: nothing needs clearing
  hipStream_t s = rt().stream;
  k_courant<<<nb, kDiagThreads, 0, s>>>(g, du, dv, dw, ddx, ddy, ddz, *dt, ws + 8);
  k_courant_final<<<1, 1024, 0, s>>>(ws + 8, nb, ws);
  VK_LAUNCH_CHECK("check_courant_noalloc");
  return fetch(ws, courant, 1, s);
}

int vk_monitor(const int* nx, const int* ny, const int* nz, const int* l, const double* t,
               const double* dx, const double* dy, const double* dz, double* dif,
               double* t_all, double* t_dif, const double* t_before, const double* u,
               const double* v, const double* w, double* vel_dif, double* u_dif,
               double* v_dif, double* w_dif, double* u_int_before, double* v_int_before,
               double* w_int_before, double* t_b, const double* dt, const int* int_report,
               int* loc) {
  VK_REQUIRE_INIT();
  DIAG_GRID();
  DIAG_FIELD(dtt, t);
  DIAG_FIELD(dtb, t_before);
  DIAG_FIELD(du, u);
  DIAG_FIELD(dv, v);
  DIAG_FIELD(dw, w);
  VK_RESOLVE(ddx, const double, dx, (g.nx + 2) * sizeof(double));
  VK_RESOLVE(ddy, const double, dy, (g.ny + 2) * sizeof(double));
  VK_RESOLVE(ddz, const double, dz, (g.nz + 2) * sizeof(double));
  if (!dif || !t_all || !t_dif || !vel_dif || !u_dif || !v_dif || !w_dif || !u_int_before ||
      !v_int_before || !w_int_before || !t_b)
    return fail(VK_ERR_ARGUMENT, "monitor: a scalar output is null");
  const int nb = diag_blocks(g);
  double* ws = static_cast<double*>(workspace(kWorkDiag, (size_t)(6 * kDiagBlocks + 8) *
                                                             sizeof(double)));
  if (!ws) return VK_ERR_HIP;
  hipStream_t s = rt().stream;
  {
    ProfScope ps("monitor", s);
    k_monitor<<<nb, kDiagThreads, 0, s>>>(g, du, dv, dw, dtt, dtb, ddx, ddy, ddz, ws + 8);
    k_monitor_final<<<1, 1024, 0, s>>>(g, ws + 8, nb, ws);
  }
  VK_LAUNCH_CHECK("monitor");
  double r[6];
  int rc = fetch(ws, r, 6, s);
  if (rc) return rc;
  // the host side of nse_dominik.f:2807-2871, in the reference's order
  if (*l == 1) {
    *t_b = 0.0;
    *u_int_before = 0.0;
    *v_int_before = 0.0;
    *w_int_before = 0.0;
  }
  const double u_int = r[0], v_int = r[1], w_int = r[2];
  *u_dif = std::fabs((u_int - *u_int_before) / u_int);
  *v_dif = std::fabs((v_int - *v_int_before) / v_int);
  *w_dif = std::fabs((w_int - *w_int_before) / w_int);
  *vel_dif = (*u_dif >= *v_dif) ? *u_dif : *v_dif;  // dmax1(u_dif, v_dif): w left out
  *u_int_before = u_int;
  *v_int_before = v_int;
  *w_int_before = w_int;
  *t_all = r[3];
  int xyz[3] = {0, 0, 0};
  const long long qe = (long long)r[5];
  if (r[4] > 0.0 && qe < (long long)g.interior()) {
    xyz[0] = 1 + (int)(qe % g.nx);
    xyz[1] = 1 + (int)((qe / g.nx) % g.ny);
    xyz[2] = 1 + (int)(qe / ((long long)g.nx * g.ny));
  }
  const bool report = int_report && *int_report == 1;
  if (report) {
    std::printf(" %.17g %.17g\n T_all = %.17g\n time error of C is: %.17g\n", *t_all, *t_b,
                *t_all, (*t_all - *t_b) / *dt);
  }
  *dif = std::fabs((*t_all - *t_b) / *t_all);
  *t_b = *t_all;
  *t_dif = r[4];
  if (report) std::printf(" %d %d %d\n", xyz[0], xyz[1], xyz[2]);
  if (loc) {
    loc[0] = xyz[0];
    loc[1] = xyz[1];
    loc[2] = xyz[2];
  }
  return VK_OK;
}

int vk_check_convergence(const double* y, const double* yl, const double* u, const int* l,
                         const double* dt, const double* re, const int* nx, const int* ny,
                         const int* nz, double* error_sum) {
  VK_REQUIRE_INIT();
  (void)l;
  (void)dt;
  DIAG_GRID();
  DIAG_FIELD(du, u);
  VK_RESOLVE(dy_, const double, y, (g.ny + 2) * sizeof(double));
  if (!error_sum) return fail(VK_ERR_ARGUMENT, "check_convergence: error_sum is null");
  double* ws = static_cast<double*>(workspace(kWorkDiag, (size_t)(6 * kDiagBlocks + 8) *
                                                             sizeof(double)));
  if (!ws) return VK_ERR_HIP;
  hipStream_t s = rt().stream;
  {
    ProfScope ps("check_convergence", s);
    const int nb = g.nz < kDiagBlocks ? g.nz : kDiagBlocks;
    k_convergence<<<nb, kDiagThreads, 0, s>>>(g, du, dy_, *yl, *re, ws + 8);
    k_sum_final<<<1, 1024, 0, s>>>(ws + 8, nb, ws);
  }
  VK_LAUNCH_CHECK("check_convergence");
  return fetch(ws, error_sum, 1, s);
}

}  // extern "C"
