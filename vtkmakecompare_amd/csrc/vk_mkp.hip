// vk_mkp.hip -- the pressure projection MKP (mkp.f:9-220): right-hand side
// div(u_temp)/dt fused into the forward FFT, spectral y solves (SPN,
// SP0B), inverse FFT, halo fill, then gradient, velocity correction,
// pressure, splitting-error diagnostic, u_ta and divergence in one pass.
#include <cstring>

#include "vk_kernels.h"
#include "vk_runtime.h"

namespace vk {

// mkp.f:54-84: KXZ(0:nx/2,0:nz) and AC(1:ny,3), PI = 3.14159265358979312
__global__ void k_mkp_coefficients(Geom g, Metrics m, double xl, double zl, double* kxz,
                                   double* ac) {
  const double pi = 3.14159265358979312;
  const int nh = g.nx / 2;
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  int nkxz = (nh + 1) * (g.nz + 1);
  if (t < nkxz) {
    int I = t % (nh + 1), K = t / (nh + 1);
    double kx = (double)I / xl * pi * 2.0;
    double kz = (K <= g.nz / 2) ? (double)K : (double)(K - g.nz);
    if (K <= g.nz - 1) kz = kz / zl * pi * 2.0;
    double dxI = m.dx[I], dzK = m.dz[K];
    kxz[t] = 2.0 * (1.0 - cos(kx * dxI)) / dxI / dxI + 2.0 * (1.0 - cos(kz * dzK)) / dzK / dzK;
  } else if (t < nkxz + g.ny) {
    int J = t - nkxz + 1;
    double a1 = 1.0 / m.dy12[J - 1] / m.dy[J];
    double a3 = 1.0 / m.dy12[J] / m.dy[J];
    ac[J - 1] = a1;
    ac[J - 1 + 2 * g.ny] = a3;
    ac[J - 1 + g.ny] = -a1 - a3;
  }
}

// mkp.f:141-217 at one (i,j,k) of 0..n: gradient of PHI at the faces, the
// corrected velocity, and on the interior p = PHI, the diagnostic
// u_ep2 = -dt grad(p) + u_ta (read before u_ta is replaced; p's halo at
// nx+1, ny+1, nz+1 is the caller's, as in the reference), u_ta = dt grad,
// and divu of the corrected velocity.  Every value is computed before the
// first store, so all the point's loads can be in flight together (a
// store would otherwise fence the loads after it: the outputs may alias
// the inputs as far as the compiler knows).  INTERIOR: i, j, k >= 1 known
// at compile time; EP2: the u_ep2 diagnostic is produced.
template <bool INTERIOR, bool EP2>
__device__ inline void project_point(const Geom& g, const RMetrics& r, const double* phi,
                                     const double* ut, const double* vt, const double* wt,
                                     double dt, const MkpOut& o, int i, int j, int k) {
  long long c = g.at(i, j, k);
  const long long sy = g.sy, sz = g.sz;
  const double ph = phi[c], phx = phi[c + 1], phy = phi[c + sy], phz = phi[c + sz];
  const double gx = (phx - ph) * r.rdx12[i];
  const double gy = (phy - ph) * r.rdy12[j];
  const double gz = (phz - ph) * r.rdz12[k];
  const double un = ut[c] - dt * gx, vn = vt[c] - dt * gy, wn = wt[c] - dt * gz;
  const bool interior = INTERIOR || (i >= 1 && j >= 1 && k >= 1);
  double div = 0.0, ex = 0.0, ey = 0.0, ez = 0.0;
  if (interior) {
    if (EP2) {
      const double px = (i < g.nx) ? phx : o.p[c + 1];
      const double py = (j < g.ny) ? phy : o.p[c + sy];
      const double pz = (k < g.nz) ? phz : o.p[c + sz];
      ex = -dt * (px - ph) * r.rdx12[i] + o.u_ta[c];
      ey = -dt * (py - ph) * r.rdy12[j] + o.v_ta[c];
      ez = -dt * (pz - ph) * r.rdz12[k] + o.w_ta[c];
    }
    const double um = ut[c - 1] - dt * ((ph - phi[c - 1]) * r.rdx12[i - 1]);
    const double vm = vt[c - sy] - dt * ((ph - phi[c - sy]) * r.rdy12[j - 1]);
    const double wm = wt[c - sz] - dt * ((ph - phi[c - sz]) * r.rdz12[k - 1]);
    div = (un - um) * r.rdx[i] + (vn - vm) * r.rdy[j] + (wn - wm) * r.rdz[k];
  }
  o.u[c] = un;
  o.v[c] = vn;
  o.w[c] = wn;
  if (interior) {
    o.p[c] = ph;
    if (EP2) {
      o.u_ep2[c] = ex;
      o.v_ep2[c] = ey;
      o.w_ep2[c] = ez;
    }
    o.divu[c] = div;
  }
  o.u_ta[c] = dt * gx;
  o.v_ta[c] = dt * gy;
  o.w_ta[c] = dt * gz;
}

// the interior points, a flat launch (flat_interior)
template <bool EP2>
__global__ __launch_bounds__(kFlatThreads) void k_project(Geom g, FlatRange fr, RMetrics r,
                                                          const double* phi, const double* ut,
                                                          const double* vt, const double* wt,
                                                          double dt, MkpOut o) {
  int i, j, k;
  if (!flat_point(g, flat_cell(fr), i, j, k)) return;
  project_point<true, EP2>(g, r, phi, ut, vt, wt, dt, o, i, j, k);
}

// the points of 0..n with an index 0: the faces i = 0, j = 0, k = 0
template <typename Fn>
__device__ inline void zero_faces(const Geom& g, long long q, Fn&& fn) {
  const long long a = (long long)(g.ny + 1) * (g.nz + 1);         // i = 0
  const long long b = (long long)g.nx * (g.nz + 1);               // j = 0, i >= 1
  const long long c = (long long)g.nx * g.ny;                     // k = 0, i, j >= 1
  if (q < a) {
    fn(0, (int)(q % (g.ny + 1)), (int)(q / (g.ny + 1)));
  } else if (q < a + b) {
    q -= a;
    fn(1 + (int)(q % g.nx), 0, (int)(q / g.nx));
  } else if (q < a + b + c) {
    q -= a + b;
    fn(1 + (int)(q % g.nx), 1 + (int)(q / g.nx), 0);
  }
}
inline long long zero_face_count(const Geom& g) {
  return (long long)(g.ny + 1) * (g.nz + 1) + (long long)g.nx * (g.nz + 1) +
         (long long)g.nx * g.ny;
}

template <bool EP2>
__global__ void k_project_faces(Geom g, RMetrics r, const double* phi, const double* ut,
                                const double* vt, const double* wt, double dt, MkpOut o) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  zero_faces(g, q, [&](int i, int j, int k) {
    project_point<false, EP2>(g, r, phi, ut, vt, wt, dt, o, i, j, k);
  });
}

int launch_mkp(const Geom& g, const Metrics& m, const double* u_temp, const double* v_temp,
               const double* w_temp, double dt, double xl, double zl, const MkpOut& o,
               hipStream_t s) {
  if (int frc = fft_sizes_ok(g)) return frc;
  const int nh = g.nx / 2;
  size_t nk = (size_t)(nh + 1) * (g.nz + 1) + 3 * (size_t)g.ny;
  double* coef = static_cast<double*>(workspace(kWorkMisc, (nk + 8) * sizeof(double)));
  double* phit = static_cast<double*>(workspace(kWorkSpec, 2 * g.spec_cells() * sizeof(double)));
  double* phi = static_cast<double*>(workspace(kWorkPhi, g.cells() * sizeof(double)));
  if (!coef || !phit || !phi) return VK_ERR_HIP;
  double* kxz = coef;
  double* ac = coef + (size_t)(nh + 1) * (g.nz + 1);
  int nthreads = (int)nk;
  {
    ProfScope ps("mkp_coef", s);
    k_mkp_coefficients<<<(nthreads + 255) / 256, 256, 0, s>>>(g, m, xl, zl, kxz, ac);
  }
  RMetrics r;
  int rc = make_rmetrics(g, m, r, s);
  if (rc) return rc;
  rc = launch_p2s_div(g, r, u_temp, v_temp, w_temp, dt, phit, s);
  if (rc) return rc;
  // the zero mode's serial march (SP0B) in the spectral solve's first
  // block, beside the other modes' solves
  rc = launch_spn_sp0b(g, phit, ac, kxz, s);
  if (rc) return rc;
  rc = launch_s2p_inplace<double>(g, phit, phi, s, "cfft_z_inv", "irfft_x");
  if (rc) return rc;
  launch_halo_neumann(g, phi, s);
  ProfScope ps("project", s);
  const long long nf = zero_face_count(g);
  const unsigned nfb = (unsigned)((nf + 255) / 256);
  const FlatRange fr = flat_interior(g);
  if (o.u_ep2) {
    k_project<true><<<fr.nb, kFlatThreads, 0, s>>>(g, fr, r, phi, u_temp, v_temp, w_temp, dt, o);
    k_project_faces<true><<<nfb, 256, 0, s>>>(g, r, phi, u_temp, v_temp, w_temp, dt, o);
  } else {
    k_project<false><<<fr.nb, kFlatThreads, 0, s>>>(g, fr, r, phi, u_temp, v_temp, w_temp, dt, o);
    k_project_faces<false><<<nfb, 256, 0, s>>>(g, r, phi, u_temp, v_temp, w_temp, dt, o);
  }
  return VK_OK;
}

// Mixed-precision projection: phi = phi1 + delta (+ psi after a correction
// pass) from float Poisson solutions, summed in FP64 at every point read;
// the corrected velocity, p and divu are FP64, u_ta is stored in float.
// Same index ranges and formulas as project_point (mkp.f:141-217); the
// u_ep2 diagnostic is not produced.  Returns |divu| at the point (0 on the
// index-0 faces, where divu is not defined).  PSI: the third solution o.psi
// is added (known at compile time, so that no phi read waits on a branch
// and every load of the point can be in flight together).
template <bool INTERIOR, bool PSI>
__device__ inline double project_mixed_point(const Geom& g, const RMetrics& r, const float* phi1,
                                             const float* dphi, const double* ut,
                                             const double* vt, const double* wt, double dt,
                                             const MkpOutMixed& o, int i, int j, int k) {
  const long long c = g.at(i, j, k);
  const long long sy = g.sy, sz = g.sz;
  const float* psi = o.psi;
  auto phi = [&](long long q) {
    const double a = (double)phi1[q] + (double)dphi[q];
    return PSI ? a + (double)psi[q] : a;
  };
  const double ph = phi(c);
  const double gx = (phi(c + 1) - ph) * r.rdx12[i];
  const double gy = (phi(c + sy) - ph) * r.rdy12[j];
  const double gz = (phi(c + sz) - ph) * r.rdz12[k];
  const double un = ut[c] - dt * gx, vn = vt[c] - dt * gy, wn = wt[c] - dt * gz;
  const bool interior = INTERIOR || (i >= 1 && j >= 1 && k >= 1);
  // every value computed before the first store, as in project_point
  double div = 0.0;
  if (interior) {
    const double um = ut[c - 1] - dt * ((ph - phi(c - 1)) * r.rdx12[i - 1]);
    const double vm = vt[c - sy] - dt * ((ph - phi(c - sy)) * r.rdy12[j - 1]);
    const double wm = wt[c - sz] - dt * ((ph - phi(c - sz)) * r.rdz12[k - 1]);
    div = (un - um) * r.rdx[i] + (vn - vm) * r.rdy[j] + (wn - wm) * r.rdz[k];
  }
  o.u[c] = un;
  o.v[c] = vn;
  o.w[c] = wn;
  if (interior) {
    o.p[c] = ph;
    o.divu[c] = div;
  }
  o.u_ta[c] = (float)(dt * gx);
  o.v_ta[c] = (float)(dt * gy);
  o.w_ta[c] = (float)(dt * gz);
  return fabs(div);
}

// Slots of the post-projection |divu| maximum: each block folds its
// largest value into slot (block index mod kDivSlots), so no address takes
// more than a few hundred atomics per projection; the host takes the
// maximum over the slots.
constexpr int kDivSlots = 1024;

// the block's largest |divu| into its slot: non-negative doubles order as
// their bit patterns do, so an integer maximum over the bits is the
// floating-point maximum.  Every lane of the block must call it.
static_assert(kFlatThreads == kBX * kBY, "block_divmax folds blocks of kBX x kBY lanes");
__device__ inline void block_divmax(double a, unsigned long long* slots) {
  __shared__ unsigned long long wmax[kBX * kBY / 64];
  unsigned long long b = (unsigned long long)__double_as_longlong(a);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(b, off);
    b = o > b ? o : b;
  }
  const int t = threadIdx.x + blockDim.x * threadIdx.y;
  if ((t & 63) == 0) wmax[t >> 6] = b;
  __syncthreads();
  if (t == 0) {
    for (int q = 1; q < kBX * kBY / 64; q++) b = wmax[q] > b ? wmax[q] : b;
    const unsigned blk = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    atomicMax(&slots[blk % kDivSlots], b);
  }
}

// the interior points, a flat launch (flat_interior)
template <bool PSI>
__global__ __launch_bounds__(kFlatThreads) void k_project_mixed(Geom g, FlatRange fr, RMetrics r,
                                                                const float* phi1,
                                                                const float* dphi,
                                                                const double* ut,
                                                                const double* vt,
                                                                const double* wt, double dt,
                                                                MkpOutMixed o) {
  int i, j, k;
  // lanes on halo cells take part in the block maximum with 0
  const double a = flat_point(g, flat_cell(fr), i, j, k)
                       ? project_mixed_point<true, PSI>(g, r, phi1, dphi, ut, vt, wt, dt, o, i,
                                                        j, k)
                       : 0.0;
  block_divmax(a, o.divmax);
}

__global__ void k_clear_slots(unsigned long long* p) {
  p[blockIdx.x * blockDim.x + threadIdx.x] = 0ull;
}

// The slots' maximum into *out, a pinned host word the host reads once the
// stream has drained, and the slots back to 0 for the next projection (one
// block of kDivSlots threads).
__global__ __launch_bounds__(kDivSlots) void k_divmax_final(unsigned long long* slots,
                                                            unsigned long long* out) {
  __shared__ unsigned long long wm[kDivSlots / 64];
  const int t = threadIdx.x;
  unsigned long long b = slots[t];
  slots[t] = 0ull;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(b, off);
    b = o > b ? o : b;
  }
  if ((t & 63) == 0) wm[t >> 6] = b;
  __syncthreads();
  if (t == 0) {
    for (int q = 1; q < kDivSlots / 64; q++) b = wm[q] > b ? wm[q] : b;
    *out = b;
  }
}

namespace {
// one pinned, device-visible host word per process for the readback, and
// its device address
unsigned long long* g_divmax_host = nullptr;
unsigned long long* g_divmax_dev = nullptr;
void release_divmax_host() {
  if (g_divmax_host) (void)hipHostFree(g_divmax_host);
  g_divmax_host = nullptr;
  g_divmax_dev = nullptr;
}
bool divmax_word(unsigned long long** host, unsigned long long** dev) {
  on_finalize(release_divmax_host);
  if (!g_divmax_host) {
    if (hipHostMalloc((void**)&g_divmax_host, 64, hipHostMallocCoherent | hipHostMallocMapped) !=
        hipSuccess) {
      g_divmax_host = nullptr;
      fail(VK_ERR_HIP, "mkp: pinned readback word");
      return false;
    }
    if (hipHostGetDevicePointer((void**)&g_divmax_dev, g_divmax_host, 0) != hipSuccess) {
      (void)hipHostFree(g_divmax_host);
      g_divmax_host = nullptr;
      g_divmax_dev = nullptr;
      fail(VK_ERR_HIP, "mkp: device address of the readback word");
      return false;
    }
  }
  *host = g_divmax_host;
  *dev = g_divmax_dev;
  return true;
}
}  // namespace

template <bool PSI>
__global__ void k_project_mixed_faces(Geom g, RMetrics r, const float* phi1, const float* dphi,
                                      const double* ut, const double* vt, const double* wt,
                                      double dt, MkpOutMixed o) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  zero_faces(g, q, [&](int i, int j, int k) {
    project_mixed_point<false, PSI>(g, r, phi1, dphi, ut, vt, wt, dt, o, i, j, k);
  });
}

// The projection of the mixed-precision step.  Pass 1 solves
// L phi1 = div(u_temp)/dt with float transforms and spectra (the right-hand
// side is formed, and kept, in FP64; every tridiagonal elimination, SP0B's
// march of the zero mode included, runs in FP64 arithmetic).  Pass 2 is
// one iterative-refinement step: the residual div(u_temp)/dt - L phi1 is
// formed in FP64 from the kept right-hand side and the stored phi1, and
// solved the same way for delta.  The projection uses phi1 + delta in FP64
// and reports the largest |divu|.  Where that is above kMixedDivTol (the
// first steps from a state far from solenoidal, or while the penalisation
// still cuts the velocity sharply at the solid's surface), the
// refinement continues once more: the residual is the post-projection
// divergence itself, div(u)/dt from the FP64 velocity, solved the same way
// for psi, and the projection is redone from u_temp with phi1 + delta +
// psi.  Deciding that reads one 8-byte word back to the host, so the host
// waits for the stream here.
int launch_mkp_mixed(const Geom& g, const Metrics& m, const double* u_temp,
                     const double* v_temp, const double* w_temp, double dt, double xl, double zl,
                     const MkpOutMixed& o, hipStream_t s) {
  if (int frc = fft_sizes_ok(g)) return frc;
  const int nh = g.nx / 2;
  size_t nk = (size_t)(nh + 1) * (g.nz + 1) + 3 * (size_t)g.ny;
  double* coef = static_cast<double*>(workspace(kWorkMisc, (nk + 8) * sizeof(double)));
  float* phit = static_cast<float*>(workspace(kWorkSpec, 2 * g.spec_cells() * sizeof(float)));
  float* phi1 = static_cast<float*>(workspace(kWorkPhi, g.cells() * sizeof(float)));
  float* dphi = static_cast<float*>(workspace(kWorkPhi2, g.cells() * sizeof(float)));
  double* rhs = static_cast<double*>(workspace(kWorkRhs, g.cells() * sizeof(double)));
  if (!coef || !phit || !phi1 || !dphi || !rhs) return VK_ERR_HIP;
  double* kxz = coef;
  double* ac = coef + (size_t)(nh + 1) * (g.nz + 1);
  {
    ProfScope ps("mkp_coef", s);
    k_mkp_coefficients<<<((int)nk + 255) / 256, 256, 0, s>>>(g, m, xl, zl, kxz, ac);
  }
  RMetrics r;
  int rc = make_rmetrics(g, m, r, s);
  if (rc) return rc;
  // pass 1
  rc = launch_p2s_div_store(g, r, u_temp, v_temp, w_temp, dt, rhs, phit, s);
  if (rc) return rc;
  rc = launch_spn_sp0b<float>(g, phit, ac, kxz, s);
  if (rc) return rc;
  rc = launch_s2p_inplace<float>(g, phit, phi1, s, "cfft_z_inv", "irfft_x");
  if (rc) return rc;
  launch_halo_neumann<float>(g, phi1, s);
  // pass 2: refinement on the FP64 residual
  rc = launch_p2s_resid(g, r, rhs, phi1, phit, s);
  if (rc) return rc;
  rc = launch_spn_sp0b<float>(g, phit, ac, kxz, s);
  if (rc) return rc;
  rc = launch_s2p_inplace<float>(g, phit, dphi, s, "cfft_z_inv2", "irfft_x2");
  if (rc) return rc;
  launch_halo_neumann<float>(g, dphi, s);
  bool fresh_slots = false;
  unsigned long long* divmax = static_cast<unsigned long long*>(
      workspace(kWorkDivMax, kDivSlots * sizeof(unsigned long long), &fresh_slots));
  unsigned long long *hmax = nullptr, *dmaxp = nullptr;
  if (!divmax || !divmax_word(&hmax, &dmaxp)) return VK_ERR_HIP;
  // k_divmax_final leaves the slots at 0 for the next projection; a new
  // buffer is cleared here
  if (fresh_slots) k_clear_slots<<<kDivSlots / 256, 256, 0, s>>>(divmax);
  MkpOutMixed op = o;
  op.divmax = divmax;
  op.psi = nullptr;
  const long long nf = zero_face_count(g);
  const FlatRange fr = flat_interior(g);
  auto project = [&](const char* name) {
    ProfScope ps(name, s);
    const unsigned nfb = (unsigned)((nf + 255) / 256);
    if (op.psi) {
      k_project_mixed<true><<<fr.nb, kFlatThreads, 0, s>>>(g, fr, r, phi1, dphi, u_temp, v_temp,
                                                           w_temp, dt, op);
      k_project_mixed_faces<true><<<nfb, 256, 0, s>>>(g, r, phi1, dphi, u_temp, v_temp, w_temp,
                                                       dt, op);
    } else {
      k_project_mixed<false><<<fr.nb, kFlatThreads, 0, s>>>(g, fr, r, phi1, dphi, u_temp, v_temp,
                                                            w_temp, dt, op);
      k_project_mixed_faces<false><<<nfb, 256, 0, s>>>(g, r, phi1, dphi, u_temp, v_temp, w_temp,
                                                        dt, op);
    }
    k_divmax_final<<<1, kDivSlots, 0, s>>>(divmax, dmaxp);
  };
  project("project_mixed");
  rc = hip_check(hipStreamSynchronize(s), "mkp divmax");
  if (rc) return rc;
  const unsigned long long bits = *(volatile unsigned long long*)hmax;
  double dmax;
  std::memcpy(&dmax, &bits, sizeof(dmax));
  if (!(dmax > kMixedDivTol)) return VK_OK;
  // the correction pass: L psi = div(u)/dt of the projected velocity
  float* psi = static_cast<float*>(workspace(kWorkPhi3, g.cells() * sizeof(float)));
  if (!psi) return VK_ERR_HIP;
  rc = launch_p2s_div_f32(g, r, o.u, o.v, o.w, dt, phit, s, "rfft_x_corr", "cfft_z_corr");
  if (rc) return rc;
  rc = launch_spn_sp0b<float>(g, phit, ac, kxz, s);
  if (rc) return rc;
  rc = launch_s2p_inplace<float>(g, phit, psi, s, "cfft_z_inv_corr", "irfft_x_corr");
  if (rc) return rc;
  launch_halo_neumann<float>(g, psi, s);
  op.psi = psi;
  project("project_mixed_corr");
  return VK_OK;
}

}  // namespace vk

using namespace vk;

extern "C" {

int vk_mkp(double* p, const double* u_temp, const double* v_temp, const double* w_temp,
           const double* dt, const double* dx, const double* dy, const double* dz,
           const double* dx12, const double* dy12, const double* dz12, const double* re,
           const int* nx, const int* ny, const int* nz, const int* ompnthreads, double* u_ta,
           double* v_ta, double* w_ta, double* u_ep2, double* v_ep2, double* w_ep2,
           double* divu, double* u, double* v, double* w, const double* xl, const double* zl) {
  VK_REQUIRE_INIT();
  (void)re;
  (void)ompnthreads;
  Geom g(*nx, *ny, *nz);
  if (g.nx < 4 || g.ny < 2 || g.nz < 2) return fail(VK_ERR_BAD_SIZE, "mkp: grid too small");
  if (!g.fits32()) return fail(VK_ERR_BAD_SIZE, "mkp: grid too large");
  size_t fb = g.cells() * sizeof(double);
  VK_RESOLVE(dut, const double, u_temp, fb);
  VK_RESOLVE(dvt, const double, v_temp, fb);
  VK_RESOLVE(dwt, const double, w_temp, fb);
  MkpOut o;
  {
    VK_RESOLVE(a, double, p, fb);
    o.p = a;
  }
  {
    VK_RESOLVE(a, double, u_ta, fb);
    o.u_ta = a;
  }
  {
    VK_RESOLVE(a, double, v_ta, fb);
    o.v_ta = a;
  }
  {
    VK_RESOLVE(a, double, w_ta, fb);
    o.w_ta = a;
  }
  {
    VK_RESOLVE(a, double, u_ep2, fb);
    o.u_ep2 = a;
  }
  {
    VK_RESOLVE(a, double, v_ep2, fb);
    o.v_ep2 = a;
  }
  {
    VK_RESOLVE(a, double, w_ep2, fb);
    o.w_ep2 = a;
  }
  {
    VK_RESOLVE(a, double, divu, fb);
    o.divu = a;
  }
  {
    VK_RESOLVE(a, double, u, fb);
    o.u = a;
  }
  {
    VK_RESOLVE(a, double, v, fb);
    o.v = a;
  }
  {
    VK_RESOLVE(a, double, w, fb);
    o.w = a;
  }
  VK_RESOLVE(ddx, const double, dx, (g.nx + 2) * sizeof(double));
  VK_RESOLVE(ddy, const double, dy, (g.ny + 2) * sizeof(double));
  VK_RESOLVE(ddz, const double, dz, (g.nz + 2) * sizeof(double));
  VK_RESOLVE(ddx12, const double, dx12, (g.nx + 2) * sizeof(double));
  VK_RESOLVE(ddy12, const double, dy12, (g.ny + 2) * sizeof(double));
  VK_RESOLVE(ddz12, const double, dz12, (g.nz + 2) * sizeof(double));
  Metrics m{ddx, ddy, ddz, ddx12, ddy12, ddz12};
  int rc = launch_mkp(g, m, dut, dvt, dwt, *dt, *xl, *zl, o, rt().stream);
  if (rc) return rc;
  VK_LAUNCH_CHECK("mkp");
  return VK_OK;
}

int vk_phibc(const int* nx, const int* ny, const int* nz, double* phi) {
  VK_REQUIRE_INIT();
  Geom g(*nx, *ny, *nz);
  VK_RESOLVE(dp, double, phi, g.cells() * sizeof(double));
  launch_halo_neumann(g, dp, rt().stream);
  VK_LAUNCH_CHECK("phibc");
  return VK_OK;
}

}  // extern "C"
