// vk_tridiag.hip -- entry points of the implicit line solves: solve_u_temp
// and solve_t (ADI sweeps y -> z -> x for u, v, w; y -> x -> z for T), SPN
// and SP0B of the Poisson solve; the line solver's shared queue and
// reference-line tables; SP0B's march.  The line systems and the sweep
// launchers are in vk_linesys.h, instantiated in vk_sweep_*.hip.
#include <map>
#include <string>

#include "vk_kernels.h"
#include "vk_partition.h"
#include "vk_runtime.h"

namespace vk {

// the fast kernel's queue: two counters, each on its own 128-byte line,
// then the list; cleared when allocated, afterwards by the general kernel
namespace {
// one queue per stream (sweeps on different streams run concurrently),
// each with two counters used by alternate queued sweeps on that stream:
// the general kernel of one sweep clears the other's, which the next
// sweep then fills
struct QueueBuf {
  unsigned* p = nullptr;
  size_t cap = 0;
  unsigned parity = 0;
};
std::map<hipStream_t, QueueBuf> g_queues;

struct TableEntry {
  double* p;
  unsigned long long epoch;
};
std::map<std::string, TableEntry> g_tables;

// vk_finalize hook: the queues and tables are device memory of this
// device, and their keys (stream handles, metric addresses) mean nothing
// after the streams are destroyed
void release_line_caches() {
  for (auto& kv : g_queues)
    if (kv.second.p) (void)dev_free(kv.second.p);
  g_queues.clear();
  for (auto& kv : g_tables)
    if (kv.second.p) (void)dev_free(kv.second.p);
  g_tables.clear();
}
}  // namespace

LineQueue line_queue(unsigned nlines, hipStream_t s) {
  on_finalize(release_line_caches);
  QueueBuf& q = g_queues[s];
  if (q.cap < nlines || !q.p) {
    if (q.p) {
      (void)hipStreamSynchronize(s);
      (void)dev_free(q.p);
      q.p = nullptr;
      q.cap = 0;
    }
    void* p = nullptr;
    if (hip_check(dev_malloc(&p, sizeof(unsigned) * (64 + (size_t)nlines), "line queue"),
                  "line queue")) {
      return LineQueue{nullptr, nullptr, nullptr};
    }
    q.p = static_cast<unsigned*>(p);
    q.cap = nlines;
    if (hip_check(hipMemsetAsync(q.p, 0, 64 * sizeof(unsigned), s), "line queue"))
      return LineQueue{nullptr, nullptr, nullptr};
  }
  q.parity ^= 1u;
  return LineQueue{q.p + 32 * q.parity, q.p + 32 * (q.parity ^ 1u), q.p + 64};
}

// reference-line tables, created on first use of a key; *validated says
// whether the table was already checked in this entry-point call
double* line_table(const void* key, size_t keylen, long total, bool* fresh, bool* validated) {
  on_finalize(release_line_caches);
  auto& cache = g_tables;
  std::string k(static_cast<const char*>(key), keylen);
  auto it = cache.find(k);
  *fresh = false;
  const unsigned long long now = call_epoch();
  if (it != cache.end()) {
    *validated = it->second.epoch == now;
    it->second.epoch = now;
    return it->second.p;
  }
  double* p = nullptr;
  if (dev_malloc(reinterpret_cast<void**>(&p), sizeof(double) * (size_t)total,
                 "reference-line table") != hipSuccess)
    return nullptr;
  cache[k] = TableEntry{p, now};
  *fresh = true;
  *validated = false;
  return p;
}

// SP0B (mkp.f:263-290): the zero mode by the forward march (sp0b_march)
// in a one-block kernel.
template <typename TS>
__global__ void k_sp0b(Geom g, const double* ac, TS* phit) {
  extern __shared__ double sh[];  // 2*(ny+2) doubles of the mode, 3 ny of AC
  sp0b_march<TS>(g, ac, phit, sh, sh + 2 * (g.ny + 2));
}

template <typename TS>
void launch_sp0b(const Geom& g, TS* phit, const double* ac, hipStream_t s) {
  ProfScope ps("sp0b", s);
  k_sp0b<TS><<<1, 256, (2 * (g.ny + 2) + 3 * g.ny) * sizeof(double), s>>>(g, ac, phit);
}
template void launch_sp0b<double>(const Geom&, double*, const double*, hipStream_t);
template void launch_sp0b<float>(const Geom&, float*, const double*, hipStream_t);

}  // namespace vk

using namespace vk;

extern "C" {

int vk_solve_u_temp(double* u_a, double* v_a, double* w_a, const int* nx, const int* ny,
                    const int* nz, const double* dx, const double* dy, const double* dz,
                    const double* dx12, const double* dy12, const double* dz12, const double* re,
                    const double* dt, const double* eta, const double* cellvalue,
                    const double* u, const double* v, const double* w) {
  VK_REQUIRE_INIT();
  Geom g(*nx, *ny, *nz);
  if (g.nx < 4 || g.ny < 4 || g.nz < 4) return fail(VK_ERR_BAD_SIZE, "solve_u_temp: n >= 4");
  if (!g.fits32()) return fail(VK_ERR_BAD_SIZE, "solve_u_temp: grid too large");
  size_t fb = g.cells() * sizeof(double);
  VK_RESOLVE(dua, double, u_a, fb);
  VK_RESOLVE(dva, double, v_a, fb);
  VK_RESOLVE(dwa, double, w_a, fb);
  VK_RESOLVE(du, const double, u, fb);
  VK_RESOLVE(dv, const double, v, fb);
  VK_RESOLVE(dw, const double, w, fb);
  VK_RESOLVE(dcv, const double, cellvalue, g.interior() * 4 * sizeof(double));
  VK_RESOLVE(ddx, const double, dx, (g.nx + 2) * sizeof(double));
  VK_RESOLVE(ddy, const double, dy, (g.ny + 2) * sizeof(double));
  VK_RESOLVE(ddz, const double, dz, (g.nz + 2) * sizeof(double));
  VK_RESOLVE(ddx12, const double, dx12, (g.nx + 2) * sizeof(double));
  VK_RESOLVE(ddy12, const double, dy12, (g.ny + 2) * sizeof(double));
  VK_RESOLVE(ddz12, const double, dz12, (g.nz + 2) * sizeof(double));
  Metrics m{ddx, ddy, ddz, ddx12, ddy12, ddz12};
  hipStream_t s = rt().stream;
  RMetrics r;
  int rrc = make_rmetrics(g, m, r, s);
  if (rrc) return rrc;
  double beta = *dt / *re / 2.0, dteta = *dt * *eta;
  double* comps[3] = {dua, dva, dwa};
  const double* olds[3] = {du, dv, dw};
  for (int c = 0; c < 3; c++) {
    launch_pen_divide(g, comps[c], olds[c], dcv, c + 2, *eta, *dt, g.ny, s);
    int rc = launch_adi_velocity(g, m, r, comps[c], dcv, c, beta, dteta, s);
    if (rc) return rc;
  }
  VK_LAUNCH_CHECK("solve_u_temp");
  return VK_OK;
}

int vk_solve_t(double* u_t, const int* nx, const int* ny, const int* nz, const double* dx,
               const double* dy, const double* dz, const double* dx12, const double* dy12,
               const double* dz12, const double* re, const double* sc, const double* dt,
               const double* rhocp, const double* thc, const double* cellvalue,
               const double* eta) {
  VK_REQUIRE_INIT();
  Geom g(*nx, *ny, *nz);
  if (g.nx < 4 || g.ny < 4 || g.nz < 4) return fail(VK_ERR_BAD_SIZE, "solve_t: n >= 4");
  if (!g.fits32()) return fail(VK_ERR_BAD_SIZE, "solve_t: grid too large");
  size_t fb = g.cells() * sizeof(double);
  VK_RESOLVE(dut, double, u_t, fb);
  VK_RESOLVE(drc, const double, rhocp, fb);
  VK_RESOLVE(dk, const double, thc, fb);
  VK_RESOLVE(dcv, const double, cellvalue, g.interior() * 4 * sizeof(double));
  VK_RESOLVE(ddx, const double, dx, (g.nx + 2) * sizeof(double));
  VK_RESOLVE(ddy, const double, dy, (g.ny + 2) * sizeof(double));
  VK_RESOLVE(ddz, const double, dz, (g.nz + 2) * sizeof(double));
  VK_RESOLVE(ddx12, const double, dx12, (g.nx + 2) * sizeof(double));
  VK_RESOLVE(ddy12, const double, dy12, (g.ny + 2) * sizeof(double));
  VK_RESOLVE(ddz12, const double, dz12, (g.nz + 2) * sizeof(double));
  Metrics m{ddx, ddy, ddz, ddx12, ddy12, ddz12};
  RMetrics r;
  int rrc = make_rmetrics(g, m, r, rt().stream);
  if (rrc) return rrc;
  double beta = *dt / *re / *sc / 2.0, dteta = *dt * *eta;
  int lrc = launch_adi_temperature(g, m, r, dut, dcv, drc, dk, beta, dteta, rt().stream);
  if (lrc) return lrc;
  VK_LAUNCH_CHECK("solve_t");
  return VK_OK;
}

int vk_spn(const int* nx, const int* ny, const int* nz, const double* ac, const double* kxz,
           double* phit) {
  VK_REQUIRE_INIT();
  Geom g(*nx, *ny, *nz);
  if (g.nx % 2 || g.ny < 2) return fail(VK_ERR_BAD_SIZE, "spn: nx even, ny >= 2");
  if (!g.fits32()) return fail(VK_ERR_BAD_SIZE, "spn: grid too large");
  VK_RESOLVE(dac, const double, ac, 3 * g.ny * sizeof(double));
  VK_RESOLVE(dk, const double, kxz, (g.nx / 2 + 1) * (g.nz + 1) * sizeof(double));
  VK_RESOLVE(dp, double, phit, 2 * g.spec_cells() * sizeof(double));
  int lrc = launch_spn(g, dp, dac, dk, rt().stream);
  if (lrc) return lrc;
  VK_LAUNCH_CHECK("spn");
  return VK_OK;
}

int vk_sp0b(const int* nx, const int* ny, const int* nz, const double* ac, double* phit) {
  VK_REQUIRE_INIT();
  Geom g(*nx, *ny, *nz);
  VK_RESOLVE(dac, const double, ac, 3 * g.ny * sizeof(double));
  VK_RESOLVE(dp, double, phit, 2 * g.spec_cells() * sizeof(double));
  launch_sp0b(g, dp, dac, rt().stream);
  VK_LAUNCH_CHECK("sp0b");
  return VK_OK;
}

}  // extern "C"
