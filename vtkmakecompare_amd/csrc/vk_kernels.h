// vk_kernels.h -- launchers shared between the translation units of the
// library (all enqueue on the given stream; none synchronises).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "vk_common.h"

namespace vk {

// halos (vk_stencil.hip); TT / T = double or float
template <typename TT>
void launch_set_bc(const Geom& g, double* u, double* v, double* w, double* p, TT* t,
                   hipStream_t s);
template <typename T>
void launch_halo_neumann(const Geom& g, T* f, hipStream_t s);
void launch_convect(const Geom& g, const Metrics& m, const double* u, const double* v,
                    const double* w, double* uv, double* vv, double* wv, double dt,
                    hipStream_t s);
void launch_pen_divide(const Geom& g, double* x, const double* old, const double* cv, int n,
                       double eta, double dt, int jmax, hipStream_t s);
template <typename T>
void launch_cv_faces(const Geom& g, T* cv, hipStream_t s);

// Dirty-line masks of one step, for the line sweeps' recorded-multiplier
// path: bit 0 (temperature: cellvalue slot 1 non-zero, rhoCp or thc not 1)
// and bits 1-3 (u, v, w: cellvalue slot 2-4 non-zero) of the word of every
// line holding such a point.  fx[(j-1) + ny (k-1)] are the x lines,
// fy[(i-1) + nx (k-1)] the y lines, fz[(i-1) + nx (j-1)] the z lines.  A
// clear bit lets a sweep skip the per-point test of its line's inputs; a
// set bit sends the line to the general path, so masks that mark too much
// only cost time.  Null pointers: each sweep tests its points itself.
struct LineFlags {
  unsigned *fx, *fy, *fz;
};
// the masks in workspace, cleared on the stream (vk_step.hip)
int line_flags_clear(const Geom& g, LineFlags& lf, hipStream_t s);

// Sets bits mk of interior point (i, j, k) in the masks of its three
// lines.  Called by every active lane of an interior block (64 x 4): the
// lanes of a wave then hold consecutive i of one (j, k), and the x line's
// word takes a single atomic per wave.
static_assert(kBX == 64, "a wave of an interior block must cover one x row");
__device__ inline void mark_dirty_lines(const LineFlags& lf, const Geom& g, int i, int j, int k,
                                        unsigned mk) {
  unsigned mx = 0;
#pragma unroll
  for (int b = 0; b < 4; b++)
    if (__ballot((mk >> b) & 1u)) mx |= 1u << b;
  if (mk) {
    atomicOr(&lf.fy[(i - 1) + g.nx * (k - 1)], mk);
    atomicOr(&lf.fz[(i - 1) + g.nx * (j - 1)], mk);
  }
  const unsigned long long act = __ballot(1);
  if (mx && (int)__lane_id() == __ffsll(act) - 1) atomicOr(&lf.fx[(j - 1) + g.ny * (k - 1)], mx);
}

// the mask bits of a point from its cellvalue slots 1-4, rhoCp and thc
template <typename T>
__device__ inline unsigned dirty_bits(T cv1, T cv2, T cv3, T cv4, T rc, T thc) {
  return (((cv1 != T(0)) | (rc != T(1)) | (thc != T(1))) ? 1u : 0u) | (cv2 != T(0) ? 2u : 0u) |
         (cv3 != T(0) ? 4u : 0u) | (cv4 != T(0) ? 8u : 0u);
}

// line solves (vk_tridiag.hip); return non-zero for unsupported lengths
// The spectrum may be stored in double or float (TS); the velocity
// systems take cellvalue in double or float (TC) with the field in double;
// the temperature systems store field, cellvalue, rhoCp and thc in TT.
// The elimination itself is FP64 arithmetic in every case.
template <typename TS>
int launch_spn(const Geom& g, TS* phit, const double* ac, const double* kxz, hipStream_t s);
template <typename TS>
void launch_sp0b(const Geom& g, TS* phit, const double* ac, hipStream_t s);
// SPN and SP0B together (the zero mode's march in the solve's first block
// where it fits, else beside it on an auxiliary stream)
template <typename TS>
int launch_spn_sp0b(const Geom& g, TS* phit, const double* ac, const double* kxz, hipStream_t s);

// SP0B (mkp.f:263-290): the zero mode (I, K) = (0, 0) of the spectrum by
// the forward march with phi(0) = phi(1) = 0.  Every thread of the calling
// block takes part: the block loads the mode into smode (2 (ny + 2)
// doubles) and the AC coefficients into sac (3 ny doubles) in parallel --
// the march would otherwise wait on a global load at every row -- one lane
// marches, and the block stores the mode.
template <typename TS>
__device__ inline void sp0b_march(const Geom& g, const double* ac, TS* phit, double* smode,
                                  double* sac) {
  const int ny = g.ny;
  double* sh = smode;
  for (int J = threadIdx.x; J <= ny + 1; J += blockDim.x) {
    long long o = 2 * g.spec_at(0, J, 0);
    sh[2 * J] = phit[o];
    sh[2 * J + 1] = phit[o + 1];
  }
  for (int q = threadIdx.x; q < 3 * ny; q += blockDim.x) sac[q] = ac[q];
  __syncthreads();
  if (threadIdx.x == 0) {
    double prevr = 0.0, previ = 0.0, curr = 0.0, curi = 0.0;  // phi(J-1), phi(J)
    for (int J = 1; J <= ny - 1; J++) {
      double a1 = sac[J - 1], a2 = sac[J - 1 + ny], a3 = sac[J - 1 + 2 * ny];
      double dr = sh[2 * J], di = sh[2 * J + 1];
      double nr = (dr - a1 * prevr - a2 * curr) / a3;
      double ni = (di - a1 * previ - a2 * curi) / a3;
      sh[2 * J] = curr;  // reuse: slot J now holds phi(J)
      sh[2 * J + 1] = curi;
      prevr = curr;
      previ = curi;
      curr = nr;
      curi = ni;
    }
    sh[2 * ny] = curr;
    sh[2 * ny + 1] = curi;
    sh[0] = 0.0;
    sh[1] = 0.0;
    sh[2 * (ny + 1)] = curr;
    sh[2 * (ny + 1) + 1] = curi;
  }
  __syncthreads();
  for (int J = threadIdx.x; J <= ny + 1; J += blockDim.x) {
    long long o = 2 * g.spec_at(0, J, 0);
    phit[o] = static_cast<TS>(sh[2 * J]);
    phit[o + 1] = static_cast<TS>(sh[2 * J + 1]);
  }
}
template <typename TC>
int launch_adi_velocity(const Geom& g, const Metrics& m, const RMetrics& r, double* f,
                        const TC* cv, int comp, double beta, double dteta, hipStream_t s,
                        const LineFlags* fl = nullptr);
template <typename TT>
int launch_adi_temperature(const Geom& g, const Metrics& m, const RMetrics& r, TT* f,
                           const TT* cv, const TT* rc, const TT* thc, double beta, double dteta,
                           hipStream_t s, TT* out = nullptr, const LineFlags* fl = nullptr);
// one sweep: velocity dir 0, 1, 2 = y, z, x; temperature dir 0, 1, 2 = y, x, z
template <typename TC>
int launch_velocity_sweep(const Geom& g, const Metrics& m, const RMetrics& r, double* f,
                          const TC* cv, int comp, int dir, double beta, double dteta,
                          hipStream_t s, const LineFlags* fl = nullptr);
template <typename TT>
int launch_temperature_sweep(const Geom& g, const Metrics& m, const RMetrics& r, TT* f,
                             const TT* cv, const TT* rc, const TT* thc, int dir, double beta,
                             double dteta, hipStream_t s, TT* out = nullptr,
                             const LineFlags* fl = nullptr);
// the transposed sweeps (vk_sweep_adj.hip), in place on the adjoint field
// (fl: the dirty-line masks of the forward step being transposed)
int launch_velocity_sweep_adj(const Geom& g, const Metrics& m, const RMetrics& r, double* lam,
                              const double* cv, int comp, int dir, double beta, double dteta,
                              hipStream_t s, const LineFlags* fl = nullptr);
int launch_temperature_sweep_adj(const Geom& g, const Metrics& m, const RMetrics& r,
                                 double* lam, const double* cv, const double* rc,
                                 const double* thc, int dir, double beta, double dteta,
                                 hipStream_t s, const LineFlags* fl = nullptr);
// reciprocal metric arrays in workspace (recomputed each call, 6(n+2) doubles)
int make_rmetrics(const Geom& g, const Metrics& m, RMetrics& r, hipStream_t s);

// FFT (vk_fft.hip); return a VK_ status
int launch_p2s(const Geom& g, const double* fp, double* phit, hipStream_t s);
// R = double or float: the transform's arithmetic and the spectrum's and
// output's storage; the divergence and residual samples are FP64.
template <typename R>
int launch_p2s_div(const Geom& g, const RMetrics& r, const double* u, const double* v,
                   const double* w, double dt, R* phit, hipStream_t s);
// the same into a float spectrum, timed under the given names
int launch_p2s_div_f32(const Geom& g, const RMetrics& r, const double* u, const double* v,
                       const double* w, double dt, float* phit, hipStream_t s, const char* xname,
                       const char* zname);
// float spectrum of div(u_temp)/dt, which is also stored in FP64 in `rhs`
int launch_p2s_div_store(const Geom& g, const RMetrics& r, const double* u, const double* v,
                         const double* w, double dt, double* rhs, float* phit, hipStream_t s);
// float spectrum of rhs - L phi for a float phi with filled halos
int launch_p2s_resid(const Geom& g, const RMetrics& r, const double* rhs, const float* phi,
                     float* phit, hipStream_t s);
template <typename R>
int launch_s2p_inplace(const Geom& g, R* phit, R* fp, hipStream_t s, const char* zname,
                       const char* xname);

// VK_OK when the spectral Poisson solve supports the grid (vk_fft.hip)
int fft_sizes_ok(const Geom& g);

// projection (vk_mkp.hip)
struct MkpOut {
  double *p, *u_ta, *v_ta, *w_ta, *u_ep2, *v_ep2, *w_ep2, *divu, *u, *v, *w;
};
int launch_mkp(const Geom& g, const Metrics& m, const double* u_temp, const double* v_temp,
               const double* w_temp, double dt, double xl, double zl, const MkpOut& o,
               hipStream_t s);

// the mixed-precision projection: FP64 velocity, p and divu, float u_ta
struct MkpOutMixed {
  double* p;
  float *u_ta, *v_ta, *w_ta;
  double *divu, *u, *v, *w;
  // set by launch_mkp_mixed: slots of the largest |divu| of the projection
  // (bits of non-negative doubles, atomicMax) and the correction pass's
  // solution
  unsigned long long* divmax = nullptr;
  const float* psi = nullptr;
};
// post-projection divergence above which launch_mkp_mixed adds a further
// correction pass: just under the 1e-10 bar the mixed form is held to, so
// every step ends below the bar.  (After one refinement the divergence is
// about 1e-11 at 384x193x384 and 4-6e-11 at 768x385x768, where a trigger
// at half the bar ran the correction on most steps.)
constexpr double kMixedDivTol = 9e-11;
int launch_mkp_mixed(const Geom& g, const Metrics& m, const double* u_temp,
                     const double* v_temp, const double* w_temp, double dt, double xl, double zl,
                     const MkpOutMixed& o, hipStream_t s);

// explicit phase of the FP64 fused step (vk_step.hip): all explicit terms,
// penalisation, AB2 and both divides; writes u_temp.. (ua, va, wa), the
// temperature work field (ut) and the history arrays in place
struct ExplicitArgs {
  const double *u, *v, *w, *t, *rc, *thc, *u_ta, *v_ta, *w_ta, *cv;
  double *u_pre, *v_pre, *w_pre, *t_pre;
  double *ua, *va, *wa, *ut;
  double dt, re, sc, eta, a1, a2, b1, c0, cd;
  LineFlags lf;  // filled when non-null
};
// the scalar part of ExplicitArgs for step l (AB2 weights, CN factors)
ExplicitArgs explicit_args(double dt, double re, double sc, double eta, int l);
// the explicit phase of u, v, w and T (one kernel, the velocity read once)
void launch_explicit_fused(const Geom& g, const Metrics& m, const RMetrics& r,
                           const ExplicitArgs& x, hipStream_t s);
// the spectrum (double) of f(i,j,k) / dy(j), index 0 of x and z read as
// the periodic image n (for the transposed Poisson solve)
int launch_p2s_wscaled(const Geom& g, const double* f, const double* rdy, double* phit,
                       hipStream_t s);
}  // namespace vk
