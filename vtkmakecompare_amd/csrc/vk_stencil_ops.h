// vk_stencil_ops.h -- point formulas of the explicit terms, one device
// function per reference expression.  They keep the reference's operation
// order (the library is built with -ffp-contract=off), so on identical
// inputs they return the reference's bits; divisions are IEEE-correct.
#pragma once

#include "vk_common.h"

namespace vk {

template <typename R>
struct F3 {  // read-only view of a halo field
  const R* __restrict__ p;
  Geom g;
  __device__ R operator()(int i, int j, int k) const { return p[g.at(i, j, k)]; }
};

// SET_CONVECT u (nse_dominik.f:774-785)
__device__ inline double conv_u(const F3<double>& U, const F3<double>& V, const F3<double>& W,
                                const Metrics& m, int i, int j, int k, double dt) {
  double c = U(i, j, k);
  double up = U(i + 1, j, k), um = U(i - 1, j, k);
  double u1 = ((c + up) * (c + up) - (c + um) * (c + um)) / m.dx12[i];
  double u2 = ((V(i + 1, j, k) * m.dx[i + 1] + V(i, j, k) * m.dx[i]) * (U(i, j + 1, k) + c) -
               (V(i + 1, j - 1, k) * m.dx[i + 1] + V(i, j - 1, k) * m.dx[i]) *
                   (c + U(i, j - 1, k))) /
              m.dx12[i] / m.dy[j];
  double u3 = ((W(i + 1, j, k) * m.dx[i + 1] + W(i, j, k) * m.dx[i]) * (U(i, j, k + 1) + c) -
               (W(i + 1, j, k - 1) * m.dx[i + 1] + W(i, j, k - 1) * m.dx[i]) *
                   (c + U(i, j, k - 1))) /
              m.dx12[i] / m.dz[k];
  return 0.25 * (-u1 - u2 - u3) * dt + 1.0 * dt;
}

// SET_CONVECT v (nse_dominik.f:794-805)
__device__ inline double conv_v(const F3<double>& U, const F3<double>& V, const F3<double>& W,
                                const Metrics& m, int i, int j, int k, double dt) {
  double c = V(i, j, k);
  double v1 = ((U(i, j + 1, k) * m.dy[j + 1] + U(i, j, k) * m.dy[j]) * (V(i + 1, j, k) + c) -
               (U(i - 1, j + 1, k) * m.dy[j + 1] + U(i - 1, j, k) * m.dy[j]) *
                   (c + V(i - 1, j, k))) /
              m.dy12[j] / m.dx[i];
  double vp = V(i, j + 1, k), vm = V(i, j - 1, k);
  double v2 = ((c + vp) * (c + vp) - (c + vm) * (c + vm)) / m.dy12[j];
  double v3 = ((W(i, j + 1, k) * m.dy[j + 1] + W(i, j, k) * m.dy[j]) * (V(i, j, k + 1) + c) -
               (W(i, j + 1, k - 1) * m.dy[j + 1] + W(i, j, k - 1) * m.dy[j]) *
                   (c + V(i, j, k - 1))) /
              m.dy12[j] / m.dz[k];
  return 0.25 * (-v1 - v2 - v3) * dt;
}

// SET_CONVECT w (nse_dominik.f:814-825)
__device__ inline double conv_w(const F3<double>& U, const F3<double>& V, const F3<double>& W,
                                const Metrics& m, int i, int j, int k, double dt) {
  double c = W(i, j, k);
  double w1 = ((U(i, j, k + 1) * m.dz[k + 1] + U(i, j, k) * m.dz[k]) * (W(i + 1, j, k) + c) -
               (U(i - 1, j, k + 1) * m.dz[k + 1] + U(i - 1, j, k) * m.dz[k]) *
                   (c + W(i - 1, j, k))) /
              m.dz12[k] / m.dx[i];
  double w2 = ((V(i, j, k + 1) * m.dz[k + 1] + V(i, j, k) * m.dz[k]) * (W(i, j + 1, k) + c) -
               (V(i, j - 1, k + 1) * m.dz[k + 1] + V(i, j - 1, k) * m.dz[k]) *
                   (c + W(i, j - 1, k))) /
              m.dz12[k] / m.dy[j];
  double wp = W(i, j, k + 1), wm = W(i, j, k - 1);
  double w3 = ((c + wp) * (c + wp) - (c + wm) * (c + wm)) / m.dz12[k];
  return 0.25 * (-w1 - w2 - w3) * dt;
}

// SET_CONVECTION_FOR_SCALAR (nse_dominik.f:852-879); `sgn` = -1 gives the
// forward form (-UT-VT-WT), +1 the adjoint.f:149-176 form (UT+VT+WT) with
// the adjoint velocity in place of u.
__device__ inline double conv_scalar(const F3<double>& U, const F3<double>& V,
                                     const F3<double>& W, const F3<double>& T,
                                     const F3<double>& RC, const Metrics& m, int i, int j, int k,
                                     double dt, bool adjoint_form) {
  double rc = RC(i, j, k), tc = T(i, j, k);
  double ut = (-((RC(i - 1, j, k) * T(i - 1, j, k) * m.dx[i] + rc * tc * m.dx[i - 1]) *
                 U(i - 1, j, k)) /
                   2.0 / m.dx12[i - 1] +
               ((rc * tc * m.dx[i + 1] + RC(i + 1, j, k) * T(i + 1, j, k) * m.dx[i]) *
                U(i, j, k)) /
                   2.0 / m.dx12[i]) /
              m.dx[i];
  double vt = (-((RC(i, j - 1, k) * T(i, j - 1, k) * m.dy[j] + rc * tc * m.dy[j - 1]) *
                 V(i, j - 1, k)) /
                   2.0 / m.dy12[j - 1] +
               ((rc * tc * m.dy[j + 1] + RC(i, j + 1, k) * T(i, j + 1, k) * m.dy[j]) *
                V(i, j, k)) /
                   2.0 / m.dy12[j]) /
              m.dy[j];
  double wt = (-((RC(i, j, k - 1) * T(i, j, k - 1) * m.dz[k] + rc * tc * m.dz[k - 1]) *
                 W(i, j, k - 1)) /
                   2.0 / m.dz12[k - 1] +
               ((rc * tc * m.dz[k + 1] + RC(i, j, k + 1) * T(i, j, k + 1) * m.dz[k]) *
                W(i, j, k)) /
                   2.0 / m.dz12[k]) /
              m.dz[k];
  if (adjoint_form) return (ut + vt + wt) * dt / rc + 1.0 * dt;
  return (-ut - vt - wt) * dt / rc + 1.0 * dt;
}

// SET_CONDUCTION (nse_dominik.f:903-939); c0 = DT/RE/2.0
__device__ inline double cond_u(const F3<double>& U, const Metrics& m, int i, int j, int k,
                                double c0) {
  double c = U(i, j, k);
  double u1 = ((U(i + 1, j, k) - c) / m.dx[i + 1] - (c - U(i - 1, j, k)) / m.dx[i]) / m.dx12[i];
  double u2 =
      ((U(i, j + 1, k) - c) / m.dy12[j] - (c - U(i, j - 1, k)) / m.dy12[j - 1]) / m.dy[j];
  double u3 =
      ((U(i, j, k + 1) - c) / m.dz12[k] - (c - U(i, j, k - 1)) / m.dz12[k - 1]) / m.dz[k];
  return c0 * (u1 + u2 + u3);
}
__device__ inline double cond_v(const F3<double>& V, const Metrics& m, int i, int j, int k,
                                double c0) {
  double c = V(i, j, k);
  double v1 =
      ((V(i + 1, j, k) - c) / m.dx12[i] - (c - V(i - 1, j, k)) / m.dx12[i - 1]) / m.dx[i];
  double v2 = ((V(i, j + 1, k) - c) / m.dy[j + 1] - (c - V(i, j - 1, k)) / m.dy[j]) / m.dy12[j];
  double v3 =
      ((V(i, j, k + 1) - c) / m.dz12[k] - (c - V(i, j, k - 1)) / m.dz12[k - 1]) / m.dz[k];
  return c0 * (v1 + v2 + v3);
}
__device__ inline double cond_w(const F3<double>& W, const Metrics& m, int i, int j, int k,
                                double c0) {
  double c = W(i, j, k);
  double w1 =
      ((W(i + 1, j, k) - c) / m.dx12[i] - (c - W(i - 1, j, k)) / m.dx12[i - 1]) / m.dx[i];
  double w2 =
      ((W(i, j + 1, k) - c) / m.dy12[j] - (c - W(i, j - 1, k)) / m.dy12[j - 1]) / m.dy[j];
  double w3 = ((W(i, j, k + 1) - c) / m.dz[k + 1] - (c - W(i, j, k - 1)) / m.dz[k]) / m.dz12[k];
  return c0 * (w1 + w2 + w3);
}

// (d1+d2)/(d1/k1+d2/k2), the harmonic face conductance
__device__ inline double hmean(double d1, double k1, double d2, double k2) {
  return (d1 + d2) / (d1 / k1 + d2 / k2);
}

// SET_DIFFUSION_FOR_SCALAR (nse_dominik.f:969-999)
__device__ inline double diff_scalar(const F3<double>& T, const F3<double>& K,
                                     const F3<double>& RC, const Metrics& m, int i, int j,
                                     int k, double re, double sc, double dt) {
  double tc = T(i, j, k), kc = K(i, j, k);
  double ut1 = (hmean(m.dx[i + 1], K(i + 1, j, k), m.dx[i], kc) * (T(i + 1, j, k) - tc) /
                    m.dx12[i] -
                hmean(m.dx[i - 1], K(i - 1, j, k), m.dx[i], kc) * (tc - T(i - 1, j, k)) /
                    m.dx12[i - 1]) /
               m.dx[i];
  double ut2 = (hmean(m.dy[j + 1], K(i, j + 1, k), m.dy[j], kc) * (T(i, j + 1, k) - tc) /
                    m.dy12[j] -
                hmean(m.dy[j - 1], K(i, j - 1, k), m.dy[j], kc) * (tc - T(i, j - 1, k)) /
                    m.dy12[j - 1]) /
               m.dy[j];
  double ut3 = (hmean(m.dz[k + 1], K(i, j, k + 1), m.dz[k], kc) * (T(i, j, k + 1) - tc) /
                    m.dz12[k] -
                hmean(m.dz[k - 1], K(i, j, k - 1), m.dz[k], kc) * (tc - T(i, j, k - 1)) /
                    m.dz12[k - 1]) /
               m.dz[k];
  return (ut1 + ut2 + ut3) / 2.0 / re / sc / RC(i, j, k) * dt;
}

// ---------------------------------------------------------------------
// Reciprocal forms used by the fused time steps.  Same expressions with
// every division by a metric replaced by a multiplication with its
// precomputed reciprocal, and the harmonic conductance evaluated with one
// division; results agree with the forms above to rounding (about 1e-16
// relative per term in FP64).  Templates over the arithmetic type R, the
// field views (F3 or F3as) and the metric sets: R = double with Metrics /
// RMetrics for the FP64 step, R = float with float metric copies for the
// mixed-precision step.

// a halo field of T read as R
template <typename T, typename R>
struct F3as {
  const T* __restrict__ p;
  Geom g;
  __device__ R operator()(int i, int j, int k) const { return static_cast<R>(p[g.at(i, j, k)]); }
};

// A halo field's values at the points (i0 + di, j0 + dj, k0 + dk) with
// |di|, |dj|, |dk| <= 1, loaded into registers before any arithmetic and
// read as R by the stencil formulas through F3's (i, j, k) interface.
// Loading them all first puts every load of a point in flight together;
// left to itself the compiler issues a few, waits, computes and issues the
// next few.  Only the points loaded may be read: after inlining every
// index is a constant, so the array lives in registers.
template <typename T, typename R>
struct Nb27 {
  T v[3][3][3];
  int i0, j0, k0;
  __device__ R operator()(int i, int j, int k) const {
    return static_cast<R>(v[i - i0 + 1][j - j0 + 1][k - k0 + 1]);
  }
  __device__ void load(const T* __restrict__ p, const Geom& g, long long o, int di, int dj,
                       int dk) {
    v[di + 1][dj + 1][dk + 1] = p[o + di + g.sy * dj + g.sz * dk];
  }
};

// Entries c - 1, c, c + 1 of a metric array, loaded up front like Nb27's
// points and read by absolute index.
template <typename R>
struct Met3 {
  R v[3];
  int c;
  __device__ R operator[](int q) const { return v[q - c + 1]; }
  template <typename P>
  __device__ void load(const P* a, int at) {
    c = at;
    v[0] = a[at - 1];
    v[1] = a[at];
    v[2] = a[at + 1];
  }
};

// The metric arrays and their reciprocals around one point (i, j, k), as
// one set serving both metric arguments of the stencil formulas.  Entries
// no formula reads are dead loads, which the compiler drops.
template <typename R>
struct PtMet {
  Met3<R> dx, dy, dz, dx12, dy12, dz12, rdx, rdy, rdz, rdx12, rdy12, rdz12;
  __device__ void load(const Metrics& m, const RMetrics& r, int i, int j, int k) {
    dx.load(m.dx, i);
    dy.load(m.dy, j);
    dz.load(m.dz, k);
    dx12.load(m.dx12, i);
    dy12.load(m.dy12, j);
    dz12.load(m.dz12, k);
    rdx.load(r.rdx, i);
    rdy.load(r.rdy, j);
    rdz.load(r.rdz, k);
    rdx12.load(r.rdx12, i);
    rdy12.load(r.rdy12, j);
    rdz12.load(r.rdz12, k);
  }
  // a metric set with float arrays d, 1/d and 1/d12 per axis (no d12)
  template <typename FM>
  __device__ void load_f(const FM& f, int i, int j, int k) {
    dx.load(f.dx, i);
    dy.load(f.dy, j);
    dz.load(f.dz, k);
    rdx.load(f.rdx, i);
    rdy.load(f.rdy, j);
    rdz.load(f.rdz, k);
    rdx12.load(f.rdx12, i);
    rdy12.load(f.rdy12, j);
    rdz12.load(f.rdz12, k);
  }
};

// the points of u, v and w that SET_CONVECT and SET_CONDUCTION read at one
// point (nse_dominik.f:774-825, :903-939)
template <typename T, typename R>
__device__ inline void load_velocity_nbhd(Nb27<T, R>& U, Nb27<T, R>& V, Nb27<T, R>& W,
                                          const T* u, const T* v, const T* w, const Geom& g,
                                          int i, int j, int k) {
  const long long o = g.at(i, j, k);
  U.i0 = V.i0 = W.i0 = i;
  U.j0 = V.j0 = W.j0 = j;
  U.k0 = V.k0 = W.k0 = k;
  Nb27<T, R>* f[3] = {&U, &V, &W};
  const T* p[3] = {u, v, w};
#pragma unroll
  for (int c = 0; c < 3; c++) {
    f[c]->load(p[c], g, o, 0, 0, 0);
    f[c]->load(p[c], g, o, 1, 0, 0);
    f[c]->load(p[c], g, o, -1, 0, 0);
    f[c]->load(p[c], g, o, 0, 1, 0);
    f[c]->load(p[c], g, o, 0, -1, 0);
    f[c]->load(p[c], g, o, 0, 0, 1);
    f[c]->load(p[c], g, o, 0, 0, -1);
  }
  U.load(u, g, o, -1, 1, 0);
  U.load(u, g, o, -1, 0, 1);
  V.load(v, g, o, 1, -1, 0);
  V.load(v, g, o, 0, -1, 1);
  W.load(w, g, o, 1, 0, -1);
  W.load(w, g, o, 0, 1, -1);
}

// the points of T, rhoCp and thc that SET_CONVECTION_FOR_SCALAR and
// SET_DIFFUSION_FOR_SCALAR read at one point (nse_dominik.f:852-879,
// :969-999): their 7-point stencils at (i, j, k).  The two routines also
// read each velocity component at the point and its lower neighbour along
// the component's axis, which load_velocity_nbhd's stencils hold.
template <typename TS, typename R>
__device__ inline void load_scalar_fields(Nb27<TS, R>& T, Nb27<TS, R>& RC, Nb27<TS, R>& K,
                                          const TS* t, const TS* rc, const TS* thc,
                                          const Geom& g, int i, int j, int k) {
  const long long o = g.at(i, j, k);
  T.i0 = RC.i0 = K.i0 = i;
  T.j0 = RC.j0 = K.j0 = j;
  T.k0 = RC.k0 = K.k0 = k;
  Nb27<TS, R>* f[3] = {&T, &RC, &K};
  const TS* p[3] = {t, rc, thc};
#pragma unroll
  for (int c = 0; c < 3; c++) {
    f[c]->load(p[c], g, o, 0, 0, 0);
    f[c]->load(p[c], g, o, 1, 0, 0);
    f[c]->load(p[c], g, o, -1, 0, 0);
    f[c]->load(p[c], g, o, 0, 1, 0);
    f[c]->load(p[c], g, o, 0, -1, 0);
    f[c]->load(p[c], g, o, 0, 0, 1);
    f[c]->load(p[c], g, o, 0, 0, -1);
  }
}

// SET_CONVECT u (nse_dominik.f:774-785)
template <typename R, typename VU, typename VV, typename VW, typename MT, typename RT>
__device__ inline R conv_u_r(const VU& U, const VV& V, const VW& W, const MT& m, const RT& r,
                             int i, int j, int k, R dt) {
  R c = U(i, j, k);
  R up = U(i + 1, j, k), um = U(i - 1, j, k);
  R u1 = ((c + up) * (c + up) - (c + um) * (c + um)) * r.rdx12[i];
  R u2 = ((V(i + 1, j, k) * m.dx[i + 1] + V(i, j, k) * m.dx[i]) * (U(i, j + 1, k) + c) -
          (V(i + 1, j - 1, k) * m.dx[i + 1] + V(i, j - 1, k) * m.dx[i]) * (c + U(i, j - 1, k))) *
         (r.rdx12[i] * r.rdy[j]);
  R u3 = ((W(i + 1, j, k) * m.dx[i + 1] + W(i, j, k) * m.dx[i]) * (U(i, j, k + 1) + c) -
          (W(i + 1, j, k - 1) * m.dx[i + 1] + W(i, j, k - 1) * m.dx[i]) * (c + U(i, j, k - 1))) *
         (r.rdx12[i] * r.rdz[k]);
  return R(0.25) * (-u1 - u2 - u3) * dt + dt;
}

// SET_CONVECT v (nse_dominik.f:794-805)
template <typename R, typename VU, typename VV, typename VW, typename MT, typename RT>
__device__ inline R conv_v_r(const VU& U, const VV& V, const VW& W, const MT& m, const RT& r,
                             int i, int j, int k, R dt) {
  R c = V(i, j, k);
  R v1 = ((U(i, j + 1, k) * m.dy[j + 1] + U(i, j, k) * m.dy[j]) * (V(i + 1, j, k) + c) -
          (U(i - 1, j + 1, k) * m.dy[j + 1] + U(i - 1, j, k) * m.dy[j]) * (c + V(i - 1, j, k))) *
         (r.rdy12[j] * r.rdx[i]);
  R vp = V(i, j + 1, k), vm = V(i, j - 1, k);
  R v2 = ((c + vp) * (c + vp) - (c + vm) * (c + vm)) * r.rdy12[j];
  R v3 = ((W(i, j + 1, k) * m.dy[j + 1] + W(i, j, k) * m.dy[j]) * (V(i, j, k + 1) + c) -
          (W(i, j + 1, k - 1) * m.dy[j + 1] + W(i, j, k - 1) * m.dy[j]) * (c + V(i, j, k - 1))) *
         (r.rdy12[j] * r.rdz[k]);
  return R(0.25) * (-v1 - v2 - v3) * dt;
}

// SET_CONVECT w (nse_dominik.f:814-825)
template <typename R, typename VU, typename VV, typename VW, typename MT, typename RT>
__device__ inline R conv_w_r(const VU& U, const VV& V, const VW& W, const MT& m, const RT& r,
                             int i, int j, int k, R dt) {
  R c = W(i, j, k);
  R w1 = ((U(i, j, k + 1) * m.dz[k + 1] + U(i, j, k) * m.dz[k]) * (W(i + 1, j, k) + c) -
          (U(i - 1, j, k + 1) * m.dz[k + 1] + U(i - 1, j, k) * m.dz[k]) * (c + W(i - 1, j, k))) *
         (r.rdz12[k] * r.rdx[i]);
  R w2 = ((V(i, j, k + 1) * m.dz[k + 1] + V(i, j, k) * m.dz[k]) * (W(i, j + 1, k) + c) -
          (V(i, j - 1, k + 1) * m.dz[k + 1] + V(i, j - 1, k) * m.dz[k]) * (c + W(i, j - 1, k))) *
         (r.rdz12[k] * r.rdy[j]);
  R wp = W(i, j, k + 1), wm = W(i, j, k - 1);
  R w3 = ((c + wp) * (c + wp) - (c + wm) * (c + wm)) * r.rdz12[k];
  return R(0.25) * (-w1 - w2 - w3) * dt;
}

// SET_CONVECTION_FOR_SCALAR (nse_dominik.f:852-879), rrc = 1/rhoCp(i,j,k)
template <typename R, typename VU, typename VV, typename VW, typename VT, typename VR,
          typename MT, typename RT>
__device__ inline R conv_scalar_r(const VU& U, const VV& V, const VW& W, const VT& T,
                                  const VR& RC, const MT& m, const RT& r, int i, int j, int k,
                                  R dt, R rrc) {
  R rc = RC(i, j, k), tc = T(i, j, k);
  R rt = rc * tc;
  R ut = (-((RC(i - 1, j, k) * T(i - 1, j, k) * m.dx[i] + rt * m.dx[i - 1]) * U(i - 1, j, k)) *
              (R(0.5) * r.rdx12[i - 1]) +
          ((rt * m.dx[i + 1] + RC(i + 1, j, k) * T(i + 1, j, k) * m.dx[i]) * U(i, j, k)) *
              (R(0.5) * r.rdx12[i])) *
         r.rdx[i];
  R vt = (-((RC(i, j - 1, k) * T(i, j - 1, k) * m.dy[j] + rt * m.dy[j - 1]) * V(i, j - 1, k)) *
              (R(0.5) * r.rdy12[j - 1]) +
          ((rt * m.dy[j + 1] + RC(i, j + 1, k) * T(i, j + 1, k) * m.dy[j]) * V(i, j, k)) *
              (R(0.5) * r.rdy12[j])) *
         r.rdy[j];
  R wt = (-((RC(i, j, k - 1) * T(i, j, k - 1) * m.dz[k] + rt * m.dz[k - 1]) * W(i, j, k - 1)) *
              (R(0.5) * r.rdz12[k - 1]) +
          ((rt * m.dz[k + 1] + RC(i, j, k + 1) * T(i, j, k + 1) * m.dz[k]) * W(i, j, k)) *
              (R(0.5) * r.rdz12[k])) *
         r.rdz[k];
  return (-ut - vt - wt) * dt * rrc + dt;
}

// SET_CONDUCTION (nse_dominik.f:903-939); c0 = DT/RE/2
template <typename R, typename VU, typename RT>
__device__ inline R cond_u_r(const VU& U, const RT& r, int i, int j, int k, R c0) {
  R c = U(i, j, k);
  R u1 = ((U(i + 1, j, k) - c) * r.rdx[i + 1] - (c - U(i - 1, j, k)) * r.rdx[i]) * r.rdx12[i];
  R u2 = ((U(i, j + 1, k) - c) * r.rdy12[j] - (c - U(i, j - 1, k)) * r.rdy12[j - 1]) * r.rdy[j];
  R u3 = ((U(i, j, k + 1) - c) * r.rdz12[k] - (c - U(i, j, k - 1)) * r.rdz12[k - 1]) * r.rdz[k];
  return c0 * (u1 + u2 + u3);
}
template <typename R, typename VV, typename RT>
__device__ inline R cond_v_r(const VV& V, const RT& r, int i, int j, int k, R c0) {
  R c = V(i, j, k);
  R v1 = ((V(i + 1, j, k) - c) * r.rdx12[i] - (c - V(i - 1, j, k)) * r.rdx12[i - 1]) * r.rdx[i];
  R v2 = ((V(i, j + 1, k) - c) * r.rdy[j + 1] - (c - V(i, j - 1, k)) * r.rdy[j]) * r.rdy12[j];
  R v3 = ((V(i, j, k + 1) - c) * r.rdz12[k] - (c - V(i, j, k - 1)) * r.rdz12[k - 1]) * r.rdz[k];
  return c0 * (v1 + v2 + v3);
}
template <typename R, typename VW, typename RT>
__device__ inline R cond_w_r(const VW& W, const RT& r, int i, int j, int k, R c0) {
  R c = W(i, j, k);
  R w1 = ((W(i + 1, j, k) - c) * r.rdx12[i] - (c - W(i - 1, j, k)) * r.rdx12[i - 1]) * r.rdx[i];
  R w2 = ((W(i, j + 1, k) - c) * r.rdy12[j] - (c - W(i, j - 1, k)) * r.rdy12[j - 1]) * r.rdy[j];
  R w3 = ((W(i, j, k + 1) - c) * r.rdz[k + 1] - (c - W(i, j, k - 1)) * r.rdz[k]) * r.rdz12[k];
  return c0 * (w1 + w2 + w3);
}

// (d1+d2)/(d1/k1+d2/k2) with one division
template <typename R>
__device__ inline R hmean_r(R d1, R k1, R d2, R k2) {
  return (d1 + d2) * k1 * k2 / (d1 * k2 + d2 * k1);
}

// SET_DIFFUSION_FOR_SCALAR (nse_dominik.f:969-999); cd = DT/(2 RE SC)
template <typename R, typename VT, typename VK, typename MT, typename RT>
__device__ inline R diff_scalar_r(const VT& T, const VK& K, const MT& m, const RT& r, int i,
                                  int j, int k, R cd, R rrc) {
  R tc = T(i, j, k), kc = K(i, j, k);
  R ut1 = (hmean_r<R>(m.dx[i + 1], K(i + 1, j, k), m.dx[i], kc) * (T(i + 1, j, k) - tc) *
               r.rdx12[i] -
           hmean_r<R>(m.dx[i - 1], K(i - 1, j, k), m.dx[i], kc) * (tc - T(i - 1, j, k)) *
               r.rdx12[i - 1]) *
          r.rdx[i];
  R ut2 = (hmean_r<R>(m.dy[j + 1], K(i, j + 1, k), m.dy[j], kc) * (T(i, j + 1, k) - tc) *
               r.rdy12[j] -
           hmean_r<R>(m.dy[j - 1], K(i, j - 1, k), m.dy[j], kc) * (tc - T(i, j - 1, k)) *
               r.rdy12[j - 1]) *
          r.rdy[j];
  R ut3 = (hmean_r<R>(m.dz[k + 1], K(i, j, k + 1), m.dz[k], kc) * (T(i, j, k + 1) - tc) *
               r.rdz12[k] -
           hmean_r<R>(m.dz[k - 1], K(i, j, k - 1), m.dz[k], kc) * (tc - T(i, j, k - 1)) *
               r.rdz12[k - 1]) *
          r.rdz[k];
  return (ut1 + ut2 + ut3) * cd * rrc;
}

}  // namespace vk
