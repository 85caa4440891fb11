// vk_kernels.h -- launchers shared between the translation units of the
// library (all enqueue on the given stream; none synchronises).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "vk_common.h"

namespace vk {

// halos (vk_stencil.hip)
void launch_set_bc(const Geom& g, double* u, double* v, double* w, double* p, double* t,
                   hipStream_t s);
void launch_halo_neumann(const Geom& g, double* f, hipStream_t s);
void launch_convect(const Geom& g, const Metrics& m, const double* u, const double* v,
                    const double* w, double* uv, double* vv, double* wv, double dt,
                    hipStream_t s);
void launch_pen_divide(const Geom& g, double* x, const double* old, const double* cv, int n,
                       double eta, double dt, int jmax, hipStream_t s);
void launch_cv_faces(const Geom& g, double* cv, hipStream_t s);

// line solves (vk_tridiag.hip); return non-zero for unsupported lengths
int launch_spn(const Geom& g, double* phit, const double* ac, const double* kxz, hipStream_t s);
void launch_sp0b(const Geom& g, double* phit, const double* ac, hipStream_t s);
int launch_adi_velocity(const Geom& g, const Metrics& m, const RMetrics& r, double* f,
                        const double* cv, int comp, double beta, double dteta, hipStream_t s);
int launch_adi_temperature(const Geom& g, const Metrics& m, const RMetrics& r, double* f,
                           const double* cv, const double* rc, const double* thc, double beta,
                           double dteta, hipStream_t s);
// reciprocal metric arrays in workspace (recomputed each call, 6(n+2) doubles)
int make_rmetrics(const Geom& g, const Metrics& m, RMetrics& r, hipStream_t s);

// FFT (vk_fft.hip); return a VK_ status
int launch_p2s(const Geom& g, const double* fp, double* phit, hipStream_t s);
int launch_p2s_div(const Geom& g, const Metrics& m, const double* u, const double* v,
                   const double* w, double dt, double* phit, hipStream_t s);
int launch_s2p_inplace(const Geom& g, double* phit, double* fp, hipStream_t s);

// projection (vk_mkp.hip)
struct MkpOut {
  double *p, *u_ta, *v_ta, *w_ta, *u_ep2, *v_ep2, *w_ep2, *divu, *u, *v, *w;
};
int launch_mkp(const Geom& g, const Metrics& m, const double* u_temp, const double* v_temp,
               const double* w_temp, double dt, double xl, double zl, const MkpOut& o,
               hipStream_t s);

// explicit phase of the fused step (vk_step.hip)
}  // namespace vk
