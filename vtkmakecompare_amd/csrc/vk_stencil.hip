// vk_stencil.hip -- halo fill, explicit terms, penalisation, AB2 combine
// and the adjoint operator kernels, with their C-ABI entry points.
#include <cstring>

#include "vk_kernels.h"
#include "vk_runtime.h"
#include "vk_stencil_ops.h"

namespace vk {

// ------------------------------------------------------------- halos

// SET_BC (nse_dominik.f:611-753).  The reference fills x, then y, then z
// faces of each field; every fill copies (or negates, or zeroes) one
// cell, so the final value of any halo cell is the interior cell reached
// by mapping each coordinate separately, times the product of the signs.
// One pass over the halo cells therefore gives the reference's values,
// and no cell it writes is read by another thread.
enum YMode { kReflect = 0, kZeroWall = 1, kNeumann = 2 };

__device__ inline bool map_y(int j, int ny, int mode, int& jm, double& s) {
  s = 1.0;
  jm = j;
  if (mode == kZeroWall) return !(j == 0 || j == ny);  // zero on both wall faces
  if (j == 0) {
    jm = 1;
    s = (mode == kReflect) ? -1.0 : 1.0;
  } else if (j == ny + 1) {
    jm = ny;
    s = (mode == kReflect) ? -1.0 : 1.0;
  }
  return true;
}

template <typename T>
__device__ inline void halo_cell(T* f, const Geom& g, int i, int j, int k, int mode) {
  int im = (i == 0) ? g.nx : (i == g.nx + 1) ? 1 : i;
  int km = (k == 0) ? g.nz : (k == g.nz + 1) ? 1 : k;
  int jm;
  double s;
  bool nz = map_y(j, g.ny, mode, jm, s);
  f[g.at(i, j, k)] = nz ? static_cast<T>(s * f[g.at(im, jm, km)]) : static_cast<T>(0);
}

template <typename TT>
__global__ void k_set_bc(Geom g, double* u, double* v, double* w, double* p, TT* t) {
  long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  int i, j, k;
  if (!halo_decode(g, q, true, i, j, k)) return;
  bool vonly = (j == g.ny) && i >= 1 && i <= g.nx && k >= 1 && k <= g.nz;
  if (v) halo_cell(v, g, i, j, k, kZeroWall);
  if (vonly) return;
  if (u) halo_cell(u, g, i, j, k, kReflect);
  if (w) halo_cell(w, g, i, j, k, kReflect);
  if (p) halo_cell(p, g, i, j, k, kNeumann);
  if (t) halo_cell(t, g, i, j, k, kReflect);
}

// full halo of one field with a given wall rule (PHIBC uses Neumann and
// agrees with mkp.f:293-320 on every cell PHIBC defines)
template <typename T>
__global__ void k_halo_one(Geom g, T* f, int mode) {
  long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  int i, j, k;
  if (!halo_decode(g, q, false, i, j, k)) return;
  halo_cell(f, g, i, j, k, mode);
}

template <typename TT>
void launch_set_bc(const Geom& g, double* u, double* v, double* w, double* p, TT* t,
                   hipStream_t s) {
  long long n = halo_count(g, true);
  ProfScope ps("set_bc", s);
  k_set_bc<TT><<<(unsigned)((n + 255) / 256), 256, 0, s>>>(g, u, v, w, p, t);
}
template void launch_set_bc<double>(const Geom&, double*, double*, double*, double*, double*,
                                    hipStream_t);
template void launch_set_bc<float>(const Geom&, double*, double*, double*, double*, float*,
                                   hipStream_t);

template <typename T>
void launch_halo_neumann(const Geom& g, T* f, hipStream_t s) {
  long long n = halo_count(g, false);
  ProfScope ps("halo_neumann", s);
  k_halo_one<T><<<(unsigned)((n + 255) / 256), 256, 0, s>>>(g, f, kNeumann);
}
template void launch_halo_neumann<double>(const Geom&, double*, hipStream_t);
template void launch_halo_neumann<float>(const Geom&, float*, hipStream_t);

// ----------------------------------------------------- explicit terms

#define INTERIOR_INDEX                               \
  int i = 1 + blockIdx.x * blockDim.x + threadIdx.x; \
  int j = 1 + blockIdx.y * blockDim.y + threadIdx.y; \
  int k = 1 + blockIdx.z;                            \
  if (i > g.nx || j > g.ny) return

__global__ void k_convect(Geom g, Metrics m, const double* u, const double* v, const double* w,
                          double* u_v, double* v_v, double* w_v, double dt) {
  INTERIOR_INDEX;
  F3<double> U{u, g}, V{v, g}, W{w, g};
  long long o = g.at(i, j, k);
  // the three values before the first store (a store would fence the later
  // loads: the outputs may alias the inputs as far as the compiler knows)
  const double a = conv_u(U, V, W, m, i, j, k, dt);
  const double b = conv_v(U, V, W, m, i, j, k, dt);
  const double c = conv_w(U, V, W, m, i, j, k, dt);
  u_v[o] = a;
  v_v[o] = b;
  w_v[o] = c;
}

__global__ void k_conv_scalar(Geom g, Metrics m, const double* u, const double* v,
                              const double* w, const double* t, const double* rc, double* out,
                              double dt, int adjoint_form) {
  INTERIOR_INDEX;
  F3<double> U{u, g}, V{v, g}, W{w, g}, T{t, g}, RC{rc, g};
  out[g.at(i, j, k)] = conv_scalar(U, V, W, T, RC, m, i, j, k, dt, adjoint_form != 0);
}

__global__ void k_conduction(Geom g, Metrics m, const double* u, const double* v,
                             const double* w, double* u_d, double* v_d, double* w_d,
                             double c0) {
  INTERIOR_INDEX;
  F3<double> U{u, g}, V{v, g}, W{w, g};
  long long o = g.at(i, j, k);
  const double a = cond_u(U, m, i, j, k, c0);
  const double b = cond_v(V, m, i, j, k, c0);
  const double c = cond_w(W, m, i, j, k, c0);
  u_d[o] = a;
  v_d[o] = b;
  w_d[o] = c;
}

__global__ void k_diffusion(Geom g, Metrics m, const double* t, const double* thc,
                            const double* rc, double* out, double re, double sc, double dt) {
  INTERIOR_INDEX;
  F3<double> T{t, g}, K{thc, g}, RC{rc, g};
  out[g.at(i, j, k)] = diff_scalar(T, K, RC, m, i, j, k, re, sc, dt);
}

// SET_IMMERSED_BOUNDARY l = 1 branch (nse_dominik.f:1037-1080): face
// values cv2..4 as averages of neighbouring cv1, copying cv1 at the last
// face (no periodic wrap there).
template <typename T>
__global__ void k_cv_faces(Geom g, T* cv) {
  INTERIOR_INDEX;
  T c = cv[g.cv(i, j, k, 1)];
  const T h = static_cast<T>(0.50);
  cv[g.cv(i, j, k, 2)] = (i < g.nx) ? (c + cv[g.cv(i + 1, j, k, 1)]) * h : c;
  cv[g.cv(i, j, k, 3)] = (j < g.ny) ? (c + cv[g.cv(i, j + 1, k, 1)]) * h : c;
  cv[g.cv(i, j, k, 4)] = (k < g.nz) ? (c + cv[g.cv(i, j, k + 1, 1)]) * h : c;
}

// SET_IMMERSED_BOUNDARY every step (nse_dominik.f:1086-1126)
__global__ void k_penalty(Geom g, const double* cv, const double* uta, const double* vta,
                          const double* wta, double* ul, double* vl, double* wl, double* tl,
                          double eta, double dt) {
  INTERIOR_INDEX;
  long long o = g.at(i, j, k);
  ul[o] = eta * cv[g.cv(i, j, k, 2)] * uta[o] * dt;
  vl[o] = eta * cv[g.cv(i, j, k, 3)] * vta[o] * dt;
  wl[o] = eta * cv[g.cv(i, j, k, 4)] * wta[o] * dt;
  tl[o] = 0.0;
}

// SET_EXPLICIT for one component (nse_dominik.f:1165-1197, 1275-1283): the
// combine, the partial halo copies (each copies one interior cell, so the
// owning thread writes them), and the history copy x_pre := x_v.
__global__ void k_explicit(Geom g, const double* xv, double* xpre, const double* xd,
                           double* xa, const double* xl, double a1, double a2, double b1) {
  INTERIOR_INDEX;
  long long o = g.at(i, j, k);
  double val = a1 * xv[o] + a2 * xpre[o] + b1 * xd[o] + b1 * xl[o];
  xa[o] = val;
  if (i == g.nx) xa[g.at(0, j, k)] = val;
  if (i == 1) xa[g.at(g.nx + 1, j, k)] = val;
  if (j == 1) xa[g.at(i, 0, k)] = -1.0 * val;
  if (j == g.ny) xa[g.at(i, g.ny + 1, k)] = -1.0 * val;
  if (k == g.nz) xa[g.at(i, j, 0)] = val;
  if (k == 1) xa[g.at(i, j, g.nz + 1)] = val;
}

__global__ void k_copy_interior(Geom g, const double* src, double* dst) {
  INTERIOR_INDEX;
  long long o = g.at(i, j, k);
  dst[o] = src[o];
}

// set_explicit_for_tempreture combine and history (nse_dominik.f:1310-1317,
// 1333-1339)
__global__ void k_explicit_t(Geom g, const double* uvt, const double* udt, double* ut,
                             double* tpre, const double* tl, double a1, double a2, double b1) {
  INTERIOR_INDEX;
  long long o = g.at(i, j, k);
  double vt = uvt[o];
  ut[o] = a1 * vt + a2 * tpre[o] + b1 * udt[o] + b1 * tl[o];
  tpre[o] = vt;
}
// its halo lines :1319-1330: x copies over all (j,k), then z copies over
// all (i,j) where the nz+1 plane takes the freshly written 0 plane
__global__ void k_explicit_t_halo_x(Geom g, double* ut) {
  int j = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y;
  if (j > g.ny + 1) return;
  ut[g.at(0, j, k)] = ut[g.at(g.nx, j, k)];
  ut[g.at(g.nx + 1, j, k)] = ut[g.at(1, j, k)];
}
__global__ void k_explicit_t_halo_z(Geom g, double* ut) {
  int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i > g.nx + 1) return;
  double v0 = ut[g.at(i, j, g.nz)];
  ut[g.at(i, j, 0)] = v0;
  ut[g.at(i, j, g.nz + 1)] = v0;
}

// main-loop temperature divide (nse_dominik.f:387-396); also the
// penalisation divide of solve_u_temp (nse_dominik.f:1358-1365) with
// x = (old + x) / (1 + eta cv_n dt)
__global__ void k_pen_divide(Geom g, double* x, const double* old, const double* cv, int n,
                             double eta, double dt, int jmax) {
  INTERIOR_INDEX;
  if (j > jmax) return;
  long long o = g.at(i, j, k);
  x[o] = (old[o] + x[o]) / (1.0 + eta * cv[g.cv(i, j, k, n)] * dt);
}

// ------------------------------------------------------ adjoint.f

// set_adj_convect (adjoint.f:16-124), typos as written
__global__ void k_adj_convect(Geom g, Metrics m, const double* u_, const double* v_,
                              const double* w_, const double* uf_, const double* vf_,
                              const double* wf_, double* u_v, double* v_v, double* w_v,
                              double dt) {
  INTERIOR_INDEX;
  F3<double> u{u_, g}, v{v_, g}, w{w_, g}, uf{uf_, g}, vf{vf_, g}, wf{wf_, g};
  const double *dx = m.dx, *dy = m.dy, *dz = m.dz, *dx12 = m.dx12, *dy12 = m.dy12,
               *dz12 = m.dz12;
  double u1 = -((uf(i, j, k) + uf(i + 1, j, k)) * (u(i, j, k) + u(i + 1, j, k)) -
                (uf(i, j, k) + uf(i - 1, j, k)) * (u(i, j, k) + u(i - 1, j, k))) /
                  dx12[i] -
              uf(i, j, k) * (u(i - 1, j, k) - u(i + 1, j, k)) / dx12[i] * 2.0;
  double u2 = -((vf(i + 1, j, k) * dx[i + 1] + vf(i, j, k) * dx[i]) *
                    (u(i, j + 1, k) + u(i, j, k)) -
                (vf(i + 1, j - 1, k) * dx[i + 1] + vf(i, j - 1, k) * dx[i]) *
                    (u(i, j, k) + u(i, j - 1, k))) /
                  dx12[i] / dy[j] -
              ((vf(i + 1, j, k) + vf(i + 1, j - 1, k)) * dx[i + 1] +
               (vf(i, j, k) + vf(i, j - 1, k)) * dx[i]) /
                  dx12[i] *
                  (-v(i + 1, j, k) - v(i + 1, j - 1, k) + v(i, j, k) + v(i, j - 1, k)) /
                  dx12[i] / 2.0;
  double u3 = -((wf(i + 1, j, k) * dx[i + 1] + wf(i, j, k) * dx[i]) *
                    (u(i, j, k + 1) + u(i, j, k)) -
                (wf(i + 1, j, k - 1) * dx[i + 1] + wf(i, j, k - 1) * dx[i]) *
                    (u(i, j, k) + u(i, j, k - 1))) /
                  dx12[i] / dz[k] -
              ((wf(i + 1, j, k) + wf(i + 1, j, k - 1)) * dx[i + 1] +
               (wf(i, j, k) + wf(i, j, k - 1)) * dx[i]) /
                  dx12[i] *
                  (-w(i + 1, j, k) - w(i + 1, j, k - 1) + w(i, j, k) + v(i, j, k - 1)) /
                  dx12[i] / 2.0;
  long long o = g.at(i, j, k);

  double v1 = ((uf(i, j + 1, k) * dy[j + 1] + uf(i, j, k) * dy[j]) *
                   (v(i + 1, j, k) + v(i, j, k)) -
               (uf(i - 1, j + 1, k) * dy[j + 1] + uf(i - 1, j, k) * dy[j]) *
                   (v(i, j, k) + v(i - 1, j, k))) /
                  dy12[j] / dx[i] -
              ((uf(i, j + 1, k) + uf(i - 1, j + 1, k)) * dy[j + 1] +
               (uf(i, j, k) + uf(i - 1, j, k)) * dy[j]) /
                  dy12[j] *
                  (-u(i, j + 1, k) - u(i - 1, j + 1, k) + u(i, j, k) + u(i - 1, j, k)) /
                  dy12[j] / 2.0;
  double v2 = ((vf(i, j, k) + vf(i, j + 1, k)) * (v(i, j, k) + v(i, j + 1, k)) -
               (vf(i, j, k) + vf(i, j - 1, k)) * (v(i, j, k) + v(i, j - 1, k))) /
                  dy12[j] -
              vf(i, j, k) * (v(i - 1, j, k) - v(i + 1, j, k)) * 2.0 / dy12[j];
  double v3 = ((wf(i, j + 1, k) * dy[j + 1] + wf(i, j, k) * dy[j]) *
                   (v(i, j, k + 1) + v(i, j, k)) -
               (wf(i, j + 1, k - 1) * dy[j + 1] + wf(i, j, k - 1) * dy[j]) *
                   (v(i, j, k) + v(i, j, k - 1))) /
                  dy12[j] / dz[k] -
              ((wf(i, j, k + 1) + wf(i, j + 1, k - 1)) * dy[j + 1] +
               (wf(i, j, k) + wf(i, j, k - 1)) * dy[j]) /
                  dy12[j] *
                  (-w(i, j + 1, k) - w(i, j + 1, k - 1) + w(i, j, k) + w(i, j, k - 1)) /
                  dy12[j] / 2.0;

  double w1 = ((uf(i, j, k + 1) * dz[k + 1] + uf(i, j, k) * dz[k]) *
                   (w(i + 1, j, k) + w(i, j, k)) -
               (uf(i - 1, j, k + 1) * dz[k + 1] + uf(i - 1, j, k) * dz[k]) *
                   (w(i, j, k) + w(i - 1, j, k))) /
                  dz12[k] / dx[i] -
              ((uf(i, j, k + 1) + uf(i - 1, j, k + 1)) * dz[k + 1] +
               (uf(i, j, k) + uf(i - 1, j, k)) * dz[k]) /
                  dz12[k] *
                  (-u(i, j, k + 1) - u(i - 1, j, k + 1) + u(i, j, k) + u(i - 1, j, k)) /
                  dz12[k] * 2.0;
  double w2 = ((vf(i, j, k + 1) * dz[k + 1] + vf(i, j, k) * dz[k]) *
                   (w(i, j + 1, k) + w(i, j, k)) -
               (vf(i, j - 1, k + 1) * dz[k + 1] + vf(i, j - 1, k) * dz[k]) *
                   (w(i, j, k) + w(i, j - 1, k))) /
                  dz12[k] / dy[j] -
              ((vf(i, j, k + 1) + vf(i, j - 1, k + 1)) * dz[k + 1] +
               (vf(i, j, k) + vf(i, j - 1, k)) * dz[k]) /
                  dz12[k] *
                  (-v(i, j, k + 1) - v(i, j - 1, k + 1) + v(i, j, k) + v(i, j - 1, k)) /
                  dz12[k] * 2.0;
  double w3 = ((wf(i, j, k) + wf(i, j, k + 1)) * (w(i, j, k) + w(i, j, k + 1)) -
               (wf(i, j, k) + wf(i, j, k - 1)) * (w(i, j, k) + w(i, j, k - 1))) /
                  dz12[k] -
              wf(i, j, k) * (w(i, j, k - 1) - w(i, j, k + 1)) * 2.0 / dz12[k];
  // stored together at the end (a store would fence the later loads)
  u_v[o] = 0.25 * (-u1 - u2 - u3) * dt + 1.0 * dt;
  v_v[o] = 0.25 * (-v1 - v2 - v3) * dt;
  w_v[o] = 0.25 * (-w1 - w2 - w3) * dt;
}

// set_adj_NS_scalar (adjoint.f:198-210) with T_final supplied
__global__ void k_adj_ns_scalar(Geom g, Metrics m, const double* t, const double* tf,
                                double* us, double* vs, double* ws, double dt) {
  INTERIOR_INDEX;
  long long o = g.at(i, j, k);
  double f = tf[o];
  us[o] = dt * f * (t[g.at(i + 1, j, k)] - t[g.at(i - 1, j, k)]) / m.dx12[i] / 2.0;
  vs[o] = dt * f * (t[g.at(i, j + 1, k)] - t[g.at(i, j - 1, k)]) / m.dy12[j] / 2.0;
  ws[o] = dt * f * (t[g.at(i, j, k + 1)] - t[g.at(i, j, k - 1)]) / m.dz12[k] / 2.0;
}

// max |f| over the interior, NaN/Inf flagged: grid-stride loop, wave and
// block reduction, one atomic per block (bit patterns of non-negative
// doubles order like the doubles)
__global__ __launch_bounds__(256) void k_absmax(Geom g, const double* f,
                                                unsigned long long* out) {
  __shared__ unsigned long long part[4];
  const long long n = (long long)g.nx * g.ny * g.nz;
  unsigned long long best = 0;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (long long)gridDim.x * blockDim.x) {
    int i = 1 + (int)(q % g.nx);
    long long r = q / g.nx;
    int j = 1 + (int)(r % g.ny), k = 1 + (int)(r / g.ny);
    double x = f[g.at(i, j, k)];
    double a = isfinite(x) ? fabs(x) : __longlong_as_double(0x7ff0000000000000LL);
    unsigned long long bits = (unsigned long long)__double_as_longlong(a);
    best = bits > best ? bits : best;
  }
  for (int off = 32; off > 0; off >>= 1) {
    unsigned long long o2 = __shfl_xor(best, off);
    best = o2 > best ? o2 : best;
  }
  if (threadIdx.x % 64 == 0) part[threadIdx.x / 64] = best;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long m = part[0];
    for (int w = 1; w < 4; w++) m = part[w] > m ? part[w] : m;
    atomicMax(out, m);
  }
}

void launch_convect(const Geom& g, const Metrics& m, const double* u, const double* v,
                    const double* w, double* uv, double* vv, double* wv, double dt,
                    hipStream_t s) {
  k_convect<<<interior_grid(g), interior_block(), 0, s>>>(g, m, u, v, w, uv, vv, wv, dt);
}

void launch_pen_divide(const Geom& g, double* x, const double* old, const double* cv, int n,
                       double eta, double dt, int jmax, hipStream_t s) {
  ProfScope ps("pen_divide", s);
  k_pen_divide<<<interior_grid(g), interior_block(), 0, s>>>(g, x, old, cv, n, eta, dt, jmax);
}

__global__ void k_rmetrics(Metrics m, int nx, int ny, int nz, double* out) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  int ax = nx + 2, ay = ny + 2, az = nz + 2;
  double* rdx = out;
  double* rdx12 = rdx + ax;
  double* rdy = rdx12 + ax;
  double* rdy12 = rdy + ay;
  double* rdz = rdy12 + ay;
  double* rdz12 = rdz + az;
  if (t < ax) {
    rdx[t] = 1.0 / m.dx[t];
    rdx12[t] = 1.0 / m.dx12[t];
  }
  if (t < ay) {
    rdy[t] = 1.0 / m.dy[t];
    rdy12[t] = 1.0 / m.dy12[t];
  }
  if (t < az) {
    rdz[t] = 1.0 / m.dz[t];
    rdz12[t] = 1.0 / m.dz12[t];
  }
}

int make_rmetrics(const Geom& g, const Metrics& m, RMetrics& r, hipStream_t s) {
  int ax = g.nx + 2, ay = g.ny + 2, az = g.nz + 2;
  double* w = static_cast<double*>(workspace(kWorkRecip, sizeof(double) * 2 * (ax + ay + az)));
  if (!w) return VK_ERR_HIP;
  // within one entry-point call the metric arrays do not change: the
  // reciprocals made earlier in the call are reused (and are not rewritten
  // while kernels on other streams read them)
  static struct {
    unsigned long long epoch = 0;
    Metrics m{};
    int nx = 0, ny = 0, nz = 0;
    double* w = nullptr;
  } made;
  const bool same = made.epoch == call_epoch() && made.w == w && made.nx == g.nx &&
                    made.ny == g.ny && made.nz == g.nz && made.m.dx == m.dx &&
                    made.m.dy == m.dy && made.m.dz == m.dz && made.m.dx12 == m.dx12 &&
                    made.m.dy12 == m.dy12 && made.m.dz12 == m.dz12;
  if (!same) {
    int nmax = ax > ay ? ax : ay;
    nmax = nmax > az ? nmax : az;
    k_rmetrics<<<(nmax + 255) / 256, 256, 0, s>>>(m, g.nx, g.ny, g.nz, w);
    made.epoch = call_epoch();
    made.m = m;
    made.nx = g.nx;
    made.ny = g.ny;
    made.nz = g.nz;
    made.w = w;
  }
  r.rdx = w;
  r.rdx12 = w + ax;
  r.rdy = w + 2 * ax;
  r.rdy12 = w + 2 * ax + ay;
  r.rdz = w + 2 * ax + 2 * ay;
  r.rdz12 = w + 2 * ax + 2 * ay + az;
  return VK_OK;
}

template <typename T>
void launch_cv_faces(const Geom& g, T* cv, hipStream_t s) {
  ProfScope ps("cv_faces", s);
  k_cv_faces<T><<<interior_grid(g), interior_block(), 0, s>>>(g, cv);
}
template void launch_cv_faces<double>(const Geom&, double*, hipStream_t);
template void launch_cv_faces<float>(const Geom&, float*, hipStream_t);

}  // namespace vk

using namespace vk;

static Metrics metrics_of(const double* dx, const double* dy, const double* dz,
                          const double* dx12, const double* dy12, const double* dz12) {
  return Metrics{dx, dy, dz, dx12, dy12, dz12};
}

#define RESOLVE_FIELD(var, ptr) VK_RESOLVE(var, double, ptr, g.cells() * sizeof(double))
#define RESOLVE_CV(var, ptr) VK_RESOLVE(var, double, ptr, g.interior() * 4 * sizeof(double))
#define RESOLVE_METRICS()                                                           \
  VK_RESOLVE(ddx, const double, dx, (g.nx + 2) * sizeof(double));                   \
  VK_RESOLVE(ddy, const double, dy, (g.ny + 2) * sizeof(double));                   \
  VK_RESOLVE(ddz, const double, dz, (g.nz + 2) * sizeof(double));                   \
  VK_RESOLVE(ddx12, const double, dx12, (g.nx + 2) * sizeof(double));               \
  VK_RESOLVE(ddy12, const double, dy12, (g.ny + 2) * sizeof(double));               \
  VK_RESOLVE(ddz12, const double, dz12, (g.nz + 2) * sizeof(double));               \
  Metrics m = metrics_of(ddx, ddy, ddz, ddx12, ddy12, ddz12)

static int check_grid(const Geom& g) {
  if (g.nx < 2 || g.ny < 2 || g.nz < 2)
    return fail(VK_ERR_BAD_SIZE, "grid must have nx, ny, nz >= 2");
  return VK_OK;
}
#define CHECK_GRID() \
  do {               \
    int _g = check_grid(g); \
    if (_g) return _g; \
  } while (0)

extern "C" {

int vk_set_bc(double* u, double* v, double* w, double* p, const int* nx, const int* ny,
              const int* nz, double* t) {
  VK_REQUIRE_INIT();
  Geom g(*nx, *ny, *nz);
  CHECK_GRID();
  RESOLVE_FIELD(du, u);
  RESOLVE_FIELD(dv, v);
  RESOLVE_FIELD(dw, w);
  RESOLVE_FIELD(dp, p);
  RESOLVE_FIELD(dt_, t);
  launch_set_bc(g, du, dv, dw, dp, dt_, rt().stream);
  VK_LAUNCH_CHECK("set_bc");
  return VK_OK;
}

int vk_set_convect(const double* u, double* u_v, const double* v, double* v_v, const double* w,
                   double* w_v, const double* dt, const int* nx, const int* ny, const int* nz,
                   const double* dx, const double* dy, const double* dz, const double* dx12,
                   const double* dy12, const double* dz12) {
  VK_REQUIRE_INIT();
  Geom g(*nx, *ny, *nz);
  CHECK_GRID();
  RESOLVE_FIELD(du, u);
  RESOLVE_FIELD(dv, v);
  RESOLVE_FIELD(dw, w);
  RESOLVE_FIELD(duv, u_v);
  RESOLVE_FIELD(dvv, v_v);
  RESOLVE_FIELD(dwv, w_v);
  RESOLVE_METRICS();
  launch_convect(g, m, du, dv, dw, duv, dvv, dwv, *dt, rt().stream);
  VK_LAUNCH_CHECK("set_convect");
  return VK_OK;
}

int vk_set_convection_for_scalar(const double* u, const double* v, const double* w,
                                 const double* dt, const int* nx, const int* ny, const int* nz,
                                 const double* dx, const double* dy, const double* dz,
                                 const double* dx12, const double* dy12, const double* dz12,
                                 const double* t, double* u_vt, const double* rhocp) {
  VK_REQUIRE_INIT();
  Geom g(*nx, *ny, *nz);
  CHECK_GRID();
  RESOLVE_FIELD(du, u);
  RESOLVE_FIELD(dv, v);
  RESOLVE_FIELD(dw, w);
  RESOLVE_FIELD(dtt, t);
  RESOLVE_FIELD(dout, u_vt);
  RESOLVE_FIELD(drc, rhocp);
  RESOLVE_METRICS();
  k_conv_scalar<<<interior_grid(g), interior_block(), 0, rt().stream>>>(g, m, du, dv, dw, dtt,
                                                                        drc, dout, *dt, 0);
  VK_LAUNCH_CHECK("set_convection_for_scalar");
  return VK_OK;
}

int vk_set_conduction(const double* u, double* u_d, const double* v, double* v_d,
                      const double* w, double* w_d, const double* re, const double* dt,
                      const int* nx, const int* ny, const int* nz, const double* dx,
                      const double* dy, const double* dz, const double* dx12,
                      const double* dy12, const double* dz12) {
  VK_REQUIRE_INIT();
  Geom g(*nx, *ny, *nz);
  CHECK_GRID();
  RESOLVE_FIELD(du, u);
  RESOLVE_FIELD(dv, v);
  RESOLVE_FIELD(dw, w);
  RESOLVE_FIELD(dud, u_d);
  RESOLVE_FIELD(dvd, v_d);
  RESOLVE_FIELD(dwd, w_d);
  RESOLVE_METRICS();
  double c0 = *dt / *re / 2.0;
  k_conduction<<<interior_grid(g), interior_block(), 0, rt().stream>>>(g, m, du, dv, dw, dud,
                                                                       dvd, dwd, c0);
  VK_LAUNCH_CHECK("set_conduction");
  return VK_OK;
}

int vk_set_diffusion_for_scalar(const double* re, const double* dt, const int* nx,
                                const int* ny, const int* nz, const double* dx,
                                const double* dy, const double* dz, const double* dx12,
                                const double* dy12, const double* dz12, const double* t,
                                double* u_dt, const double* sc, const double* thc,
                                const double* rhocp) {
  VK_REQUIRE_INIT();
  Geom g(*nx, *ny, *nz);
  CHECK_GRID();
  RESOLVE_FIELD(dtt, t);
  RESOLVE_FIELD(dout, u_dt);
  RESOLVE_FIELD(dk, thc);
  RESOLVE_FIELD(drc, rhocp);
  RESOLVE_METRICS();
  k_diffusion<<<interior_grid(g), interior_block(), 0, rt().stream>>>(g, m, dtt, dk, drc, dout,
                                                                      *re, *sc, *dt);
  VK_LAUNCH_CHECK("set_diffusion_for_scalar");
  return VK_OK;
}

int vk_set_immersed_boundary(const double* eta, const int* nx, const int* ny, const int* nz,
                             const double* dt, double* u_l, double* v_l, double* w_l,
                             const double* dx, const double* dy, const double* dz, double* t_l,
                             double* cellvalue, const double* u_ta, const double* v_ta,
                             const double* w_ta, const int* l, double* cons_cell) {
  VK_REQUIRE_INIT();
  Geom g(*nx, *ny, *nz);
  CHECK_GRID();
  (void)dx; (void)dy; (void)dz;
  RESOLVE_FIELD(dul, u_l);
  RESOLVE_FIELD(dvl, v_l);
  RESOLVE_FIELD(dwl, w_l);
  RESOLVE_FIELD(dtl, t_l);
  RESOLVE_FIELD(duta, u_ta);
  RESOLVE_FIELD(dvta, v_ta);
  RESOLVE_FIELD(dwta, w_ta);
  RESOLVE_CV(dcv, cellvalue);
  if (*l == 1) {
    if (cons_cell) *cons_cell = 0.0;
    k_cv_faces<<<interior_grid(g), interior_block(), 0, rt().stream>>>(g, dcv);
  }
  k_penalty<<<interior_grid(g), interior_block(), 0, rt().stream>>>(g, dcv, duta, dvta, dwta,
                                                                    dul, dvl, dwl, dtl, *eta,
                                                                    *dt);
  VK_LAUNCH_CHECK("set_immersed_boundary");
  return VK_OK;
}

static void ab2(int l, double& a1, double& a2, double& b1) {
  if (l == 1) {
    a1 = 1.0; a2 = 0.0; b1 = 1.0;
  } else {
    a1 = 1.5; a2 = -0.5; b1 = 1.0;
  }
}

int vk_set_explicit(const double* u_v, double* u_pre, const double* u_d, double* u_a,
                    const double* v_v, double* v_pre, const double* v_d, double* v_a,
                    const double* w_v, double* w_pre, const double* w_d, double* w_a,
                    const double* u_l, const double* v_l, const double* w_l, const int* nx,
                    const int* ny, const int* nz, const int* l) {
  VK_REQUIRE_INIT();
  Geom g(*nx, *ny, *nz);
  CHECK_GRID();
  const double* in[15] = {u_v, u_pre, u_d, u_a, v_v, v_pre, v_d, v_a, w_v, w_pre, w_d, w_a,
                          u_l, v_l, w_l};
  double* d[15];
  for (int q = 0; q < 15; q++) {
    d[q] = static_cast<double*>(resolve(in[q], g.cells() * sizeof(double), "set_explicit"));
    if (!d[q]) return last_status();
  }
  double a1, a2, b1;
  ab2(*l, a1, a2, b1);
  hipStream_t s = rt().stream;
  for (int c = 0; c < 3; c++)
    k_explicit<<<interior_grid(g), interior_block(), 0, s>>>(g, d[4 * c], d[4 * c + 1],
                                                             d[4 * c + 2], d[4 * c + 3],
                                                             d[12 + c], a1, a2, b1);
  for (int c = 0; c < 3; c++)
    k_copy_interior<<<interior_grid(g), interior_block(), 0, s>>>(g, d[4 * c], d[4 * c + 1]);
  VK_LAUNCH_CHECK("set_explicit");
  return VK_OK;
}

int vk_set_explicit_for_tempreture(const double* u_vt, const double* u_dt, double* u_t,
                                   double* t_pre, const double* t_l, const int* nx,
                                   const int* ny, const int* nz, const int* l) {
  VK_REQUIRE_INIT();
  Geom g(*nx, *ny, *nz);
  CHECK_GRID();
  RESOLVE_FIELD(dvt, u_vt);
  RESOLVE_FIELD(ddt, u_dt);
  RESOLVE_FIELD(dut, u_t);
  RESOLVE_FIELD(dtp, t_pre);
  RESOLVE_FIELD(dtl, t_l);
  double a1, a2, b1;
  ab2(*l, a1, a2, b1);
  hipStream_t s = rt().stream;
  k_explicit_t<<<interior_grid(g), interior_block(), 0, s>>>(g, dvt, ddt, dut, dtp, dtl, a1, a2,
                                                             b1);
  k_explicit_t_halo_x<<<dim3((g.ny + 2 + 63) / 64, g.nz + 2), 64, 0, s>>>(g, dut);
  k_explicit_t_halo_z<<<dim3((g.nx + 2 + 63) / 64, g.ny + 2), 64, 0, s>>>(g, dut);
  VK_LAUNCH_CHECK("set_explicit_for_tempreture");
  return VK_OK;
}

int vk_t_divide(double* u_t, const double* t, const double* cellvalue, const double* eta,
                const double* dt, const int* nx, const int* ny, const int* nz) {
  VK_REQUIRE_INIT();
  Geom g(*nx, *ny, *nz);
  CHECK_GRID();
  RESOLVE_FIELD(dut, u_t);
  RESOLVE_FIELD(dtt, t);
  RESOLVE_CV(dcv, cellvalue);
  k_pen_divide<<<interior_grid(g), interior_block(), 0, rt().stream>>>(g, dut, dtt, dcv, 1, *eta,
                                                                       *dt, g.ny);
  VK_LAUNCH_CHECK("t_divide");
  return VK_OK;
}

int vk_set_adj_convect(const double* u, double* u_v, const double* v, double* v_v,
                       const double* w, double* w_v, const double* re, const double* dt,
                       const int* nx, const int* ny, const int* nz, const double* dx,
                       const double* dy, const double* dz, const double* dx12,
                       const double* dy12, const double* dz12, const double* u_final,
                       const double* v_final, const double* w_final) {
  VK_REQUIRE_INIT();
  Geom g(*nx, *ny, *nz);
  CHECK_GRID();
  (void)re;
  RESOLVE_FIELD(du, u);
  RESOLVE_FIELD(dv, v);
  RESOLVE_FIELD(dw, w);
  RESOLVE_FIELD(duv, u_v);
  RESOLVE_FIELD(dvv, v_v);
  RESOLVE_FIELD(dwv, w_v);
  RESOLVE_FIELD(duf, u_final);
  RESOLVE_FIELD(dvf, v_final);
  RESOLVE_FIELD(dwf, w_final);
  RESOLVE_METRICS();
  k_adj_convect<<<interior_grid(g), interior_block(), 0, rt().stream>>>(
      g, m, du, dv, dw, duf, dvf, dwf, duv, dvv, dwv, *dt);
  VK_LAUNCH_CHECK("set_adj_convect");
  return VK_OK;
}

int vk_set_adj_convection_for_scalar(const double* u, const double* v, const double* w,
                                     const double* re, const double* dt, const int* nx,
                                     const int* ny, const int* nz, const double* dx,
                                     const double* dy, const double* dz, const double* dx12,
                                     const double* dy12, const double* dz12, const double* t,
                                     double* u_vt, const double* rhocp, const double* u_final,
                                     const double* v_final, const double* w_final) {
  VK_REQUIRE_INIT();
  Geom g(*nx, *ny, *nz);
  CHECK_GRID();
  (void)u; (void)v; (void)w; (void)re;
  RESOLVE_FIELD(dtt, t);
  RESOLVE_FIELD(dout, u_vt);
  RESOLVE_FIELD(drc, rhocp);
  RESOLVE_FIELD(duf, u_final);
  RESOLVE_FIELD(dvf, v_final);
  RESOLVE_FIELD(dwf, w_final);
  RESOLVE_METRICS();
  k_conv_scalar<<<interior_grid(g), interior_block(), 0, rt().stream>>>(g, m, duf, dvf, dwf, dtt,
                                                                        drc, dout, *dt, 1);
  VK_LAUNCH_CHECK("set_adj_convection_for_scalar");
  return VK_OK;
}

int vk_set_adj_ns_scalar(const double* dt, const int* nx, const int* ny, const int* nz,
                         const double* dx12, const double* dy12, const double* dz12,
                         const double* t, double* u_s, double* v_s, double* w_s,
                         const double* u_final, const double* v_final, const double* w_final,
                         const double* t_final) {
  VK_REQUIRE_INIT();
  Geom g(*nx, *ny, *nz);
  CHECK_GRID();
  (void)u_final; (void)v_final; (void)w_final;
  RESOLVE_FIELD(dtt, t);
  RESOLVE_FIELD(dtf, t_final);
  RESOLVE_FIELD(dus, u_s);
  RESOLVE_FIELD(dvs, v_s);
  RESOLVE_FIELD(dws, w_s);
  VK_RESOLVE(ddx12, const double, dx12, (g.nx + 2) * sizeof(double));
  VK_RESOLVE(ddy12, const double, dy12, (g.ny + 2) * sizeof(double));
  VK_RESOLVE(ddz12, const double, dz12, (g.nz + 2) * sizeof(double));
  Metrics m{nullptr, nullptr, nullptr, ddx12, ddy12, ddz12};
  k_adj_ns_scalar<<<interior_grid(g), interior_block(), 0, rt().stream>>>(g, m, dtt, dtf, dus,
                                                                          dvs, dws, *dt);
  VK_LAUNCH_CHECK("set_adj_ns_scalar");
  return VK_OK;
}

int vk_field_copy(const double* src, double* dst, const int* nx, const int* ny, const int* nz,
                  const int* interior_only) {
  VK_REQUIRE_INIT();
  Geom g(*nx, *ny, *nz);
  CHECK_GRID();
  RESOLVE_FIELD(ds, src);
  RESOLVE_FIELD(dd, dst);
  if (*interior_only) {
    ProfScope ps("field_copy", rt().stream);
    k_copy_interior<<<interior_grid(g), interior_block(), 0, rt().stream>>>(g, ds, dd);
    VK_LAUNCH_CHECK("field_copy");
    return VK_OK;
  }
  return hip_check(hipMemcpyAsync(dd, ds, g.cells() * sizeof(double), hipMemcpyDeviceToDevice,
                                  rt().stream),
                   "field_copy");
}

int vk_field_absmax(const double* f, const int* nx, const int* ny, const int* nz,
                    double* result) {
  VK_REQUIRE_INIT();
  Geom g(*nx, *ny, *nz);
  CHECK_GRID();
  RESOLVE_FIELD(df, f);
  unsigned long long* dres =
      static_cast<unsigned long long*>(workspace(kWorkMisc, sizeof(unsigned long long)));
  if (!dres) return VK_ERR_HIP;
  hipStream_t s = rt().stream;
  int rc = hip_check(hipMemsetAsync(dres, 0, sizeof(unsigned long long), s), "absmax memset");
  if (rc) return rc;
  k_absmax<<<1024, 256, 0, s>>>(g, df, dres);
  VK_LAUNCH_CHECK("field_absmax");
  unsigned long long bits = 0;
  rc = hip_check(hipMemcpyAsync(&bits, dres, sizeof(bits), hipMemcpyDeviceToHost, s),
                 "absmax copy");
  if (rc) return rc;
  rc = hip_check(hipStreamSynchronize(s), "absmax sync");
  if (rc) return rc;
  double r;
  std::memcpy(&r, &bits, sizeof(r));
  *result = r;
  if (!(r <= 1.7976931348623157e308)) return fail(VK_ERR_NONFINITE, "field has NaN or Inf");
  return VK_OK;
}

}  // extern "C"
