// vk_adjoint.hip -- the discrete adjoint of the FP64 fused time step and the
// backward driver: gradient of a wall-heat-flux objective with respect to
// the penalisation field cellvalue(...,1), over a forward trajectory kept
// by checkpoint-and-recompute in HBM.
//
// The reference supplies the continuous adjoint operators of adjoint.f
// (set_adj_convect, set_adj_convection_for_scalar, set_adj_NS_scalar,
// vk_stencil.hip) but no driver, objective or gradient.  The backward pass
// here is the exact transpose of the GPU forward step, so its gradient
// agrees with central finite differences of the same forward code:
//
//   forward step                     transpose in the backward pass
//   SET_BC halos                     halo adjoints folded into the cells
//                                    they copy (signs kept), halos zeroed
//   explicit phase (stencils,        pointwise weights, then for every
//   penalisation, AB2, divides)      input cell a gather over the outputs
//                                    that read it; each Jacobian entry is
//                                    the stencil itself evaluated on a
//                                    dual number seeded at that cell
//   ADI line solves A x = b          A^T lambda_b = lambda_x (transposed
//                                    tridiagonal/cyclic systems), and
//                                    dJ/dcv_r = -lambda_b,r (eta dt / alpha_r)
//                                    (x_r - b_r) from A = I - (beta/alpha) L
//   projection u = u* - dt G L^+ D u*/dt, u_ta = dt G L^+ D u*/dt
//                                    D^T, G^T as gathers; (L^+)^T from
//                                    L = W S (S symmetric, W = diag 1/dy):
//                                    W^-1 L^+ W on the nonzero (kx,kz)
//                                    modes and the transposed SP0B march
//                                    on the zero mode
//   cellvalue faces (step 1)         faces^T into slot 1
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "vk_adjoint.h"
#include "vk_kernels.h"
#include "vk_runtime.h"
#include "vk_stencil_ops.h"

namespace vk {

#define INTERIOR_POINT                               \
  int i = 1 + blockIdx.x * blockDim.x + threadIdx.x; \
  int j = 1 + blockIdx.y * blockDim.y + threadIdx.y; \
  int k = 1 + blockIdx.z;                            \
  if (i > g.nx || j > g.ny) return;                  \
  const long long o = g.at(i, j, k)

static dim3 cells_grid(const Geom& g) { return dim3((unsigned)((g.cells() + 255) / 256)); }

// ------------------------------------------------------------ halos

// SET_BC^T: every interior cell gathers the adjoints of the halo cells
// that copy it (x, z periodic; y by `mode`: 0 reflect (sign -1), 2 Neumann;
// 1 = v's rule, where the wall rows are set to zero and so pass nothing).
__global__ void k_halo_fold(Geom g, double* lam, int mode) {
  INTERIOR_POINT;
  int ic[2] = {i, -1}, kc[2] = {k, -1}, jc[2] = {j, -1};
  if (i == g.nx) ic[1] = 0;
  if (i == 1) ic[1] = g.nx + 1;
  if (k == g.nz) kc[1] = 0;
  if (k == 1) kc[1] = g.nz + 1;
  double sy = 1.0;
  if (mode != 1) {
    if (j == 1) jc[1] = 0;
    if (j == g.ny) jc[1] = g.ny + 1;
    if (g.ny == 1) jc[1] = -1;
    sy = (mode == 0) ? -1.0 : 1.0;
  }
  double acc = 0.0;
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++)
      for (int c = 0; c < 2; c++) {
        if (a + b + c == 0) continue;
        const int hi = ic[a], hj = jc[b], hk = kc[c];
        if (hi < 0 || hj < 0 || hk < 0) continue;
        acc += (b ? sy : 1.0) * lam[g.at(hi, hj, hk)];
      }
  lam[o] += acc;
}

// zero every halo cell (and, for v, the wall row j = ny that SET_BC zeroes)
__global__ void k_halo_zero(Geom g, double* lam, int vrow) {
  ALL_CELLS;
  const bool halo = !is_interior(g, qi, qj, qk);
  if (halo || (vrow && qj == g.ny)) lam[q_] = 0.0;
}

static void halo_transpose(const Geom& g, double* lam, int mode, hipStream_t s) {
  k_halo_fold<<<interior_grid(g), interior_block(), 0, s>>>(g, lam, mode);
  k_halo_zero<<<cells_grid(g), 256, 0, s>>>(g, lam, mode == 1);
}

// ------------------------------------------------------- projection^T

struct Lam3 {
  double *u, *v, *w;
};
struct CLam3 {
  const double *u, *v, *w;
};

// lambda_utemp = lambda_u on the interior; lambda_g = dt (lambda_uta - lambda_u)
__global__ void k_adj_project_in(Geom g, CLam3 lu, CLam3 luta, Lam3 lut, Lam3 lg, double dt) {
  ALL_CELLS;
  const bool in = is_interior(g, qi, qj, qk);
  // all reads before the first store (distinct buffers; a store would
  // fence the later loads otherwise)
  const double u = lu.u[q_], v = lu.v[q_], w = lu.w[q_];
  const double ut = luta.u[q_], vt = luta.v[q_], wt = luta.w[q_];
  lut.u[q_] = in ? u : 0.0;
  lut.v[q_] = in ? v : 0.0;
  lut.w[q_] = in ? w : 0.0;
  lg.u[q_] = in ? dt * (ut - u) : 0.0;
  lg.v[q_] = in ? dt * (vt - v) : 0.0;
  lg.w[q_] = in ? dt * (wt - w) : 0.0;
}

// G^T: g_x(p) = (phi(p+ex) - phi(p)) / dx12(i_p) on interior p
__global__ void k_adj_grad(Geom g, RMetrics r, CLam3 lg, double* lphi) {
  ALL_CELLS;
  double acc = 0.0;
  const bool in = is_interior(g, qi, qj, qk);
  if (in) acc -= lg.u[q_] * r.rdx12[qi] + lg.v[q_] * r.rdy12[qj] + lg.w[q_] * r.rdz12[qk];
  if (is_interior(g, qi - 1, qj, qk)) acc += lg.u[q_ - 1] * r.rdx12[qi - 1];
  if (is_interior(g, qi, qj - 1, qk)) acc += lg.v[q_ - g.sy] * r.rdy12[qj - 1];
  if (is_interior(g, qi, qj, qk - 1)) acc += lg.w[q_ - g.sz] * r.rdz12[qk - 1];
  lphi[q_] = acc;
}

// the zero mode of the transposed Poisson solve: DC(J) holds the plane
// mean of W x (W = 1/dy); m = dy DC is the plane mean of x; z = B^T m for
// SP0B's march B (phi(0) = phi(1) = 0, phi(J+1) = (D(J) - A1 phi(J-1) -
// A2 phi(J)) / A3 for J = 1..ny-1, row ny unused); DC(J) := z(J) / dy(J)
// so that the final multiplication by dy returns z.
__global__ void k_sp0b_adj(Geom g, const double* ac, const double* dy, double* phit) {
  extern __shared__ double sh[];  // ny + 2
  const int ny = g.ny;
  for (int J = threadIdx.x; J <= ny + 1; J += blockDim.x) {
    const long long o = 2 * g.spec_at(0, J, 0);
    sh[J] = (J >= 1 && J <= ny) ? phit[o] * dy[J] : 0.0;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    // lp: the adjoint of phi(J), consumed backwards; ld: that of D(J)
    double* lp = sh;
    double* ld = sh + (ny + 2);
    for (int J = 0; J <= ny + 1; J++) ld[J] = 0.0;
    for (int J = ny - 1; J >= 1; J--) {
      const double a1 = ac[J - 1], a2 = ac[J - 1 + ny], a3 = ac[J - 1 + 2 * ny];
      const double t = lp[J + 1] / a3;
      ld[J] += t;
      lp[J - 1] -= a1 * t;
      lp[J] -= a2 * t;
    }
  }
  __syncthreads();
  const double* ld = sh + (ny + 2);
  for (int J = threadIdx.x; J <= ny + 1; J += blockDim.x) {
    const long long o = 2 * g.spec_at(0, J, 0);
    phit[o] = (J >= 1 && J <= ny) ? ld[J] / dy[J] : 0.0;
    phit[o + 1] = 0.0;
  }
}

// D^T / dt with lambda_r = y dy(j): D(p) = (u(p)-u(p-ex))/dx(i) + ... on
// interior p
__global__ void k_adj_div(Geom g, RMetrics r, const double* ysol, const double* dy, double dt,
                          Lam3 lut) {
  ALL_CELLS;
  auto lr = [&](int i, int j, int k) { return ysol[g.at(i, j, k)] * dy[j] / dt; };
  double au = 0.0, av = 0.0, aw = 0.0;
  if (is_interior(g, qi, qj, qk)) {
    const double l0 = lr(qi, qj, qk);
    au += r.rdx[qi] * l0;
    av += r.rdy[qj] * l0;
    aw += r.rdz[qk] * l0;
  }
  if (is_interior(g, qi + 1, qj, qk)) au -= r.rdx[qi + 1] * lr(qi + 1, qj, qk);
  if (is_interior(g, qi, qj + 1, qk)) av -= r.rdy[qj + 1] * lr(qi, qj + 1, qk);
  if (is_interior(g, qi, qj, qk + 1)) aw -= r.rdz[qk + 1] * lr(qi, qj, qk + 1);
  lut.u[q_] += au;
  lut.v[q_] += av;
  lut.w[q_] += aw;
}

// ------------------------------------------------------------- ADI^T

// dJ/dcv += -lambda_b (eta dt / alpha) (x - b) on the rows j <= jmax
__global__ void k_adj_sweep_grad(Geom g, const double* lam, const double* x, const double* b,
                                 const double* cv, int slot, double* gcv, double dteta,
                                 int jmax) {
  INTERIOR_POINT;
  if (j > jmax) return;
  const long long c = g.cv(i, j, k, slot);
  const double alpha = 1.0 + dteta * cv[c];
  gcv[c] += -lam[o] * (dteta / alpha) * (x[o] - b[o]);
}

// ------------------------------------------------------- explicit^T

struct AdjIn {        // adjoints arriving at the explicit phase's outputs
  const double *lu, *lv, *lw, *lt;          // u_temp, v_temp, w_temp, T work
  const double *lup, *lvp, *lwp, *ltp;      // history outputs
};
struct AdjOut {       // adjoints of the step's inputs
  double *lu, *lv, *lw, *lt, *lup, *lvp, *lwp, *ltp, *luta, *lvta, *lwta;
};
__global__ void k_adj_explicit_weights(Geom g, StepIn in, AdjIn a, AdjOut out, Weights wt,
                                       double* gcv, double dt, double eta, double a1, double a2,
                                       double b1) {
  ALL_CELLS;
  if (!is_interior(g, qi, qj, qk)) {
    out.lu[q_] = 0.0; out.lv[q_] = 0.0; out.lw[q_] = 0.0; out.lt[q_] = 0.0;
    out.lup[q_] = 0.0; out.lvp[q_] = 0.0; out.lwp[q_] = 0.0; out.ltp[q_] = 0.0;
    out.luta[q_] = 0.0; out.lvta[q_] = 0.0; out.lwta[q_] = 0.0;
    return;
  }
  const long long o = q_;
  const int i = qi, j = qj, k = qk;
  const double ed = eta * dt;
  // velocity components: x_out = (x + a1 N + a2 N_pre + b1 D + b1 eta cv x_ta dt) / alpha.
  // Every input is read before the first store (a store would fence the
  // loads after it: the outputs may alias the inputs as far as the
  // compiler knows; they are distinct buffers).
  const long long c2 = g.cv(i, j, k, 2), c3 = g.cv(i, j, k, 3), c4 = g.cv(i, j, k, 4),
                  c1 = g.cv(i, j, k, 1);
  const double cu = in.cv[c2], cv_ = in.cv[c3], cw = in.cv[c4], ct = in.cv[c1];
  const double alu = a.lu[o], alv = a.lv[o], alw = a.lw[o], alt = a.lt[o];
  const double alup = a.lup[o], alvp = a.lvp[o], alwp = a.lwp[o], altp = a.ltp[o];
  const double uta = in.u_ta[o], vta = in.v_ta[o], wta = in.w_ta[o];
  const double su = in.su[o], sv = in.sv[o], sw = in.sw[o], st = in.st[o];
  const double gu = gcv[c2], gv = gcv[c3], gw = gcv[c4], gt = gcv[c1];
  const double s_u = 1.0 / (1.0 + ed * cu), mu_u = alu * s_u;
  const double s_v = 1.0 / (1.0 + ed * cv_), mu_v = alv * s_v;
  const double s_w = 1.0 / (1.0 + ed * cw), mu_w = alw * s_w;
  const double s_t = 1.0 / (1.0 + ed * ct), mu_t = alt * s_t;
  out.lu[o] = mu_u;
  out.lup[o] = a2 * mu_u;
  out.luta[o] = b1 * eta * cu * dt * mu_u;
  gcv[c2] = gu + alu * ed * s_u * (b1 * uta - su);
  wt.uv[o] = a1 * mu_u + alup;
  wt.ud[o] = b1 * mu_u;
  out.lv[o] = mu_v;
  out.lvp[o] = a2 * mu_v;
  out.lvta[o] = b1 * eta * cv_ * dt * mu_v;
  gcv[c3] = gv + alv * ed * s_v * (b1 * vta - sv);
  wt.vv[o] = a1 * mu_v + alvp;
  wt.vd[o] = b1 * mu_v;
  out.lw[o] = mu_w;
  out.lwp[o] = a2 * mu_w;
  out.lwta[o] = b1 * eta * cw * dt * mu_w;
  gcv[c4] = gw + alw * ed * s_w * (b1 * wta - sw);
  wt.wv[o] = a1 * mu_w + alwp;
  wt.wd[o] = b1 * mu_w;
  out.lt[o] = mu_t;
  out.ltp[o] = a2 * mu_t;
  gcv[c1] = gt + -alt * ed * s_t * st;
  wt.vt[o] = a1 * mu_t + altp;
  wt.dtt[o] = b1 * mu_t;
}

// ------------------------------------------------------- objective

// J = dt sum_steps sum_(i,k) dx dz [2 T(i,1,k)/dy12(0) + 2 T(i,ny,k)/dy12(ny)]:
// the wall heat flux through both walls (T's wall ghosts are the
// reflections -T), integrated over the walls and over time.
template <typename TT>
__global__ void k_objective_acc(Geom g, Metrics m, const TT* t, double dt, double* acc) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= g.nx * g.nz) return;
  const int i = 1 + q % g.nx, k = 1 + q / g.nx;
  const double w = dt * m.dx[i] * m.dz[k] * 2.0;
  acc[q] += w * ((double)t[g.at(i, 1, k)] / m.dy12[0] + (double)t[g.at(i, g.ny, k)] / m.dy12[g.ny]);
}

__global__ void k_copy_cells_f(long long n, const float* src, float* dst) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) dst[q] = src[q];
}
__global__ void k_to_double(long long n, const float* src, double* dst) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) dst[q] = (double)src[q];
}
__global__ void k_to_float(long long n, const double* src, float* dst) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) dst[q] = (float)src[q];
}

// its gradient with respect to T of one step, added to lambda_T
__global__ void k_objective_src(Geom g, Metrics m, double dt, double* lt) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= g.nx * g.nz) return;
  const int i = 1 + q % g.nx, k = 1 + q / g.nx;
  const double w = dt * m.dx[i] * m.dz[k] * 2.0;
  if (g.ny == 1) {
    lt[g.at(i, 1, k)] += w * (1.0 / m.dy12[0] + 1.0 / m.dy12[g.ny]);
    return;
  }
  lt[g.at(i, 1, k)] += w / m.dy12[0];
  lt[g.at(i, g.ny, k)] += w / m.dy12[g.ny];
}

// cellvalue faces at step 1 (SET_IMMERSED_BOUNDARY, nse_dominik.f:1037-1080),
// transposed: slot 1 gathers the face gradients
__global__ void k_adj_cv_faces(Geom g, const double* gcv, double* grad) {
  INTERIOR_POINT;
  (void)o;
  auto G = [&](int ii, int jj, int kk, int n) { return gcv[g.cv(ii, jj, kk, n)]; };
  double acc = G(i, j, k, 1);
  acc += (i < g.nx) ? 0.5 * G(i, j, k, 2) : G(i, j, k, 2);
  if (i > 1) acc += 0.5 * G(i - 1, j, k, 2);
  acc += (j < g.ny) ? 0.5 * G(i, j, k, 3) : G(i, j, k, 3);
  if (j > 1) acc += 0.5 * G(i, j - 1, k, 3);
  acc += (k < g.nz) ? 0.5 * G(i, j, k, 4) : G(i, j, k, 4);
  if (k > 1) acc += 0.5 * G(i, j, k - 1, 4);
  grad[(size_t)(i - 1) + (size_t)g.nx * ((j - 1) + (size_t)g.ny * (k - 1))] = acc;
}

__global__ void k_copy_cells(long long n, const double* src, double* dst) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) dst[q] = src[q];
}

// ------------------------------------------------------------ driver

namespace {

// The state a step carries: u, v, w, T, the four history arrays and u_ta
// (p is an output only).  All halo fields.
constexpr int kState = 11;
enum { SU, SV, SW, ST, SUP, SVP, SWP, STP, SUTA, SVTA, SWTA };

struct Dev {
  std::vector<void*> ptrs;
  ~Dev() {
    for (void* p : ptrs) (void)dev_free(p);
  }
  size_t bytes = 0;
  double* alloc(size_t n) { return static_cast<double*>(raw(n * sizeof(double))); }
  float* allocf(size_t n) { return static_cast<float*>(raw(n * sizeof(float))); }
  void* raw(size_t b) {
    void* p = nullptr;
    if (dev_malloc(&p, b, "adjoint checkpoint or tape") != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
    ptrs.push_back(p);
    bytes += b;
    return p;
  }
};

struct Problem {
  Geom g;
  Metrics m;
  double dt, re, sc, eta, xl, zl;
  const double *rc, *thc;
  double* cv;  // the run's own cellvalue (slots 2-4 written at step 1)
};

struct Work {
  double* st[kState];  // current state
  double *p, *ut, *vt, *wt, *divu;
  double* jacc;
};

hipStream_t S() { return rt().stream; }

void copy(const Geom& g, const double* src, double* dst) {
  const long long n = (long long)g.cells();
  k_copy_cells<<<(unsigned)((n + 255) / 256), 256, 0, S()>>>(n, src, dst);
}

int forward_step(const Problem& P, Work& w, int l) {
  const Geom& g = P.g;
  return vk_time_step(w.st[SU], w.st[SV], w.st[SW], w.p, w.st[ST], w.st[SUP], w.st[SVP],
                      w.st[SWP], w.st[STP], w.st[SUTA], w.st[SVTA], w.st[SWTA], w.ut, w.vt,
                      w.wt, w.divu, P.cv, P.rc, P.thc, nullptr, nullptr, nullptr, P.m.dx,
                      P.m.dy, P.m.dz, P.m.dx12, P.m.dy12, P.m.dz12, &g.nx, &g.ny, &g.nz, &P.dt,
                      &P.re, &P.sc, &P.eta, &P.xl, &P.zl, &l);
}

void objective_acc(const Problem& P, const Work& w) {
  const int n = P.g.nx * P.g.nz;
  k_objective_acc<double><<<(n + 255) / 256, 256, 0, S()>>>(P.g, P.m, w.st[ST], P.dt, w.jacc);
}

int sum_objective(const Geom& g, const double* jacc, double* out) {
  std::vector<double> h((size_t)g.nx * g.nz);
  int rc = hip_check(hipMemcpyAsync(h.data(), jacc, h.size() * sizeof(double),
                                    hipMemcpyDeviceToHost, S()),
                     "objective copy");
  if (rc) return rc;
  rc = hip_check(hipStreamSynchronize(S()), "objective sync");
  if (rc) return rc;
  double s = 0.0;
  for (double v : h) s += v;
  *out = s;
  return VK_OK;
}

// The taped step: the forward step of vk_time_step with the explicit
// outputs and every sweep's result kept.  tape[c*4 + d] for c = u, v, w, T:
// d = 0 the explicit output, 1..3 the results of the three sweeps.
// in4 receives u, v, w, T after the first SET_BC.
// lf_out: the step's dirty-line masks (in the workspace until the next
// step clears them), which the transposed sweeps of its adjoint reuse
int taped_step(const Problem& P, Work& w, int l, double* const* tape, double* const* in4,
               LineFlags* lf_out) {
  const Geom& g = P.g;
  hipStream_t s = S();
  RMetrics r;
  int rc = make_rmetrics(g, P.m, r, s);
  if (rc) return rc;
  launch_set_bc(g, w.st[SU], w.st[SV], w.st[SW], w.p, w.st[ST], s);
  if (l == 1) launch_cv_faces(g, P.cv, s);
  for (int c = 0; c < 4; c++) copy(g, w.st[c], in4[c]);
  double* twork = static_cast<double*>(workspace(kWorkGrad, g.cells() * sizeof(double)));
  if (!twork) return VK_ERR_HIP;
  ExplicitArgs x = explicit_args(P.dt, P.re, P.sc, P.eta, l);
  x.u = w.st[SU]; x.v = w.st[SV]; x.w = w.st[SW]; x.t = w.st[ST]; x.rc = P.rc; x.thc = P.thc;
  x.u_ta = w.st[SUTA]; x.v_ta = w.st[SVTA]; x.w_ta = w.st[SWTA]; x.cv = P.cv;
  x.u_pre = w.st[SUP]; x.v_pre = w.st[SVP]; x.w_pre = w.st[SWP]; x.t_pre = w.st[STP];
  x.ua = w.ut; x.va = w.vt; x.wa = w.wt; x.ut = twork;
  // the explicit kernels mark the lines next to the solid for the sweeps
  LineFlags lf{};
  rc = line_flags_clear(g, lf, s);
  if (rc) return rc;
  x.lf = lf;
  *lf_out = lf;
  launch_explicit_fused(g, P.m, r, x, s);
  const double beta_u = P.dt / P.re / 2.0, beta_t = P.dt / P.re / P.sc / 2.0;
  const double dteta = P.dt * P.eta;
  double* comps[4] = {w.ut, w.vt, w.wt, twork};
  for (int c = 0; c < 4; c++) {
    copy(g, comps[c], tape[c * 4]);
    for (int d = 0; d < 3; d++) {
      rc = (c < 3) ? launch_velocity_sweep<double>(g, P.m, r, comps[c], P.cv, c, d, beta_u,
                                                   dteta, s, &lf)
                   : launch_temperature_sweep<double>(g, P.m, r, twork, P.cv, P.rc, P.thc, d,
                                                      beta_t, dteta, s, nullptr, &lf);
      if (rc) return rc;
      copy(g, comps[c], tape[c * 4 + d + 1]);
    }
  }
  {
    // T = T work on the interior (nse_dominik.f:429-437)
    int one = 1;
    rc = vk_field_copy(twork, w.st[ST], &g.nx, &g.ny, &g.nz, &one);
    if (rc) return rc;
  }
  launch_set_bc(g, w.ut, w.vt, w.wt, w.p, w.st[ST], s);
  MkpOut o{w.p, w.st[SUTA], w.st[SVTA], w.st[SWTA], nullptr, nullptr, nullptr, w.divu,
           w.st[SU], w.st[SV], w.st[SW]};
  return launch_mkp(g, P.m, w.ut, w.vt, w.wt, P.dt, P.xl, P.zl, o, s);
}

struct AdjBufs {
  double* lam[kState];   // adjoint of the step's output state (in), then input state (out)
  double* nxt[kState];   // the other set
  double *lut[3], *lg[3], *lphi, *ysol, *spec;
  double* w8[8];
  double *coef, *gcv;
};

// the coefficient arrays of MKP (KXZ, AC) for the transposed Poisson solve
__global__ void k_adj_coefficients(Geom g, Metrics m, double xl, double zl, double* kxz,
                                   double* ac) {
  const double pi = 3.14159265358979312;
  const int nh = g.nx / 2;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int nkxz = (nh + 1) * (g.nz + 1);
  if (t < nkxz) {
    const int I = t % (nh + 1), K = t / (nh + 1);
    const double kx = (double)I / xl * pi * 2.0;
    double kz = (K <= g.nz / 2) ? (double)K : (double)(K - g.nz);
    if (K <= g.nz - 1) kz = kz / zl * pi * 2.0;
    const double dxI = m.dx[I], dzK = m.dz[K];
    kxz[t] = 2.0 * (1.0 - cos(kx * dxI)) / dxI / dxI + 2.0 * (1.0 - cos(kz * dzK)) / dzK / dzK;
  } else if (t < nkxz + g.ny) {
    const int J = t - nkxz + 1;
    const double a1 = 1.0 / m.dy12[J - 1] / m.dy[J];
    const double a3 = 1.0 / m.dy12[J] / m.dy[J];
    ac[J - 1] = a1;
    ac[J - 1 + 2 * g.ny] = a3;
    ac[J - 1 + g.ny] = -a1 - a3;
  }
}

// One step of the backward pass.  On entry B.lam holds the adjoint of the
// state after step l; on return B.lam holds the adjoint of the state
// before it, and the cellvalue gradient of step l is added to B.gcv.
int adjoint_step(const Problem& P, AdjBufs& B, int l, double* const* tape, double* const* in4,
                 const double* uta_in[3], const LineFlags* fl) {
  const Geom& g = P.g;
  hipStream_t s = S();
  RMetrics r;
  int rc = make_rmetrics(g, P.m, r, s);
  if (rc) return rc;
  const dim3 cg = cells_grid(g);
  // projection^T
  k_adj_project_in<<<cg, 256, 0, s>>>(g, CLam3{B.lam[SU], B.lam[SV], B.lam[SW]},
                                      CLam3{B.lam[SUTA], B.lam[SVTA], B.lam[SWTA]},
                                      Lam3{B.lut[0], B.lut[1], B.lut[2]},
                                      Lam3{B.lg[0], B.lg[1], B.lg[2]}, P.dt);
  k_adj_grad<<<cg, 256, 0, s>>>(g, r, CLam3{B.lg[0], B.lg[1], B.lg[2]}, B.lphi);
  halo_transpose(g, B.lphi, 2, s);
  const int nh = g.nx / 2;
  double* kxz = B.coef;
  double* ac = B.coef + (size_t)(nh + 1) * (g.nz + 1);
  {
    const int nk = (nh + 1) * (g.nz + 1) + g.ny;
    k_adj_coefficients<<<(nk + 255) / 256, 256, 0, s>>>(g, P.m, P.xl, P.zl, kxz, ac);
  }
  rc = launch_p2s_wscaled(g, B.lphi, r.rdy, B.spec, s);
  if (rc) return rc;
  rc = launch_spn<double>(g, B.spec, ac, kxz, s);
  if (rc) return rc;
  k_sp0b_adj<<<1, 256, 2 * (g.ny + 2) * sizeof(double), s>>>(g, ac, P.m.dy, B.spec);
  rc = launch_s2p_inplace<double>(g, B.spec, B.ysol, s, "cfft_z_adj", "irfft_x_adj");
  if (rc) return rc;
  k_adj_div<<<cg, 256, 0, s>>>(g, r, B.ysol, P.m.dy, P.dt, Lam3{B.lut[0], B.lut[1], B.lut[2]});
  halo_transpose(g, B.lut[0], 0, s);
  halo_transpose(g, B.lut[1], 1, s);
  halo_transpose(g, B.lut[2], 0, s);
  // ADI^T: velocity x, z, y; temperature z, x, y
  const double beta_u = P.dt / P.re / 2.0, beta_t = P.dt / P.re / P.sc / 2.0;
  const double dteta = P.dt * P.eta;
  double* lt = B.lam[ST];  // the adjoint of T after the step is that of the T work field
  for (int c = 0; c < 4; c++) {
    double* lam = (c < 3) ? B.lut[c] : lt;
    const int jmax = (c == 1) ? g.ny - 1 : g.ny;
    const int slot = (c < 3) ? c + 2 : 1;
    for (int d = 2; d >= 0; d--) {
      rc = (c < 3) ? launch_velocity_sweep_adj(g, P.m, r, lam, P.cv, c, d, beta_u, dteta, s, fl)
                   : launch_temperature_sweep_adj(g, P.m, r, lam, P.cv, P.rc, P.thc, d, beta_t,
                                                  dteta, s, fl);
      if (rc) return rc;
      k_adj_sweep_grad<<<interior_grid(g), interior_block(), 0, s>>>(
          g, lam, tape[c * 4 + d + 1], tape[c * 4 + d], P.cv, slot, B.gcv, dteta, jmax);
    }
  }
  // explicit^T
  const ExplicitArgs ea = explicit_args(P.dt, P.re, P.sc, P.eta, l);
  StepIn in{in4[0], in4[1], in4[2], in4[3], P.rc, P.thc, uta_in[0], uta_in[1], uta_in[2], P.cv,
            tape[0], tape[4], tape[8], tape[12]};
  AdjIn ai{B.lut[0], B.lut[1], B.lut[2], lt, B.lam[SUP], B.lam[SVP], B.lam[SWP], B.lam[STP]};
  AdjOut ao{B.nxt[SU],  B.nxt[SV],  B.nxt[SW],   B.nxt[ST],   B.nxt[SUP], B.nxt[SVP],
            B.nxt[SWP], B.nxt[STP], B.nxt[SUTA], B.nxt[SVTA], B.nxt[SWTA]};
  Weights wt{B.w8[0], B.w8[1], B.w8[2], B.w8[3], B.w8[4], B.w8[5], B.w8[6], B.w8[7]};
  k_adj_explicit_weights<<<cg, 256, 0, s>>>(g, in, ai, ao, wt, B.gcv, P.dt, P.eta, ea.a1,
                                            ea.a2, ea.b1);
  launch_adj_gather(0, g, P.m, r, in, wt, P.dt, ea.c0, ea.cd, B.nxt[SU], s);
  launch_adj_gather(1, g, P.m, r, in, wt, P.dt, ea.c0, ea.cd, B.nxt[SV], s);
  launch_adj_gather(2, g, P.m, r, in, wt, P.dt, ea.c0, ea.cd, B.nxt[SW], s);
  launch_adj_gather(3, g, P.m, r, in, wt, P.dt, ea.c0, ea.cd, B.nxt[ST], s);
  // SET_BC^T of the step's first halo fill
  halo_transpose(g, B.nxt[SU], 0, s);
  halo_transpose(g, B.nxt[SV], 1, s);
  halo_transpose(g, B.nxt[SW], 0, s);
  halo_transpose(g, B.nxt[ST], 0, s);
  for (int f = 0; f < kState; f++) std::swap(B.lam[f], B.nxt[f]);
  return VK_OK;
}

}  // namespace
}  // namespace vk

using namespace vk;

namespace {

struct Inputs {
  const double *u, *v, *w, *p, *t, *up, *vp, *wp, *tp, *uta, *vta, *wta, *cv, *rc, *thc;
};

int resolve_problem(Problem& P, Inputs& in, const double* u, const double* v, const double* w,
                    const double* p, const double* t, const double* u_pre, const double* v_pre,
                    const double* w_pre, const double* t_pre, const double* u_ta,
                    const double* v_ta, const double* w_ta, const double* cellvalue,
                    const double* rhocp, const double* thc, const double* dx, const double* dy,
                    const double* dz, const double* dx12, const double* dy12,
                    const double* dz12) {
  const Geom& g = P.g;
  const size_t fb = g.cells() * sizeof(double);
  const double* a[15] = {u, v, w, p, t, u_pre, v_pre, w_pre, t_pre, u_ta, v_ta, w_ta,
                         cellvalue, rhocp, thc};
  const char* names[15] = {"u", "v", "w", "p", "t", "u_pre", "v_pre", "w_pre", "t_pre",
                           "u_ta", "v_ta", "w_ta", "cellvalue", "rhocp", "thc"};
  const double* d[15];
  for (int q = 0; q < 15; q++) {
    const size_t bytes = (q == 12) ? g.interior() * 4 * sizeof(double) : fb;
    d[q] = static_cast<const double*>(resolve(a[q], bytes, names[q]));
    if (!d[q]) return last_status();
  }
  in = Inputs{d[0], d[1], d[2], d[3], d[4], d[5], d[6], d[7], d[8], d[9], d[10], d[11],
              d[12], d[13], d[14]};
  const double* mm[6] = {dx, dy, dz, dx12, dy12, dz12};
  const int len[6] = {g.nx + 2, g.ny + 2, g.nz + 2, g.nx + 2, g.ny + 2, g.nz + 2};
  const double* md[6];
  for (int q = 0; q < 6; q++) {
    md[q] = static_cast<const double*>(resolve(mm[q], len[q] * sizeof(double), "metric"));
    if (!md[q]) return last_status();
  }
  P.m = Metrics{md[0], md[1], md[2], md[3], md[4], md[5]};
  P.rc = in.rc;
  P.thc = in.thc;
  return VK_OK;
}

// allocate the work state and load the initial state into it
int init_work(const Problem& P, const Inputs& in, Dev& dev, Work& w) {
  const Geom& g = P.g;
  const size_t nc = g.cells();
  for (int f = 0; f < kState; f++)
    if (!(w.st[f] = dev.alloc(nc))) return fail(VK_ERR_HIP, "adjoint: out of device memory");
  w.p = dev.alloc(nc);
  w.ut = dev.alloc(nc);
  w.vt = dev.alloc(nc);
  w.wt = dev.alloc(nc);
  w.divu = dev.alloc(nc);
  w.jacc = dev.alloc((size_t)g.nx * g.nz);
  if (!w.p || !w.ut || !w.vt || !w.wt || !w.divu || !w.jacc)
    return fail(VK_ERR_HIP, "adjoint: out of device memory");
  const double* src[kState] = {in.u, in.v, in.w, in.t, in.up, in.vp, in.wp, in.tp,
                               in.uta, in.vta, in.wta};
  for (int f = 0; f < kState; f++) copy(g, src[f], w.st[f]);
  copy(g, in.p, w.p);
  int rc = hip_check(hipMemsetAsync(w.ut, 0, nc * sizeof(double), S()), "memset");
  rc = rc ? rc : hip_check(hipMemsetAsync(w.vt, 0, nc * sizeof(double), S()), "memset");
  rc = rc ? rc : hip_check(hipMemsetAsync(w.wt, 0, nc * sizeof(double), S()), "memset");
  rc = rc ? rc : hip_check(hipMemsetAsync(w.jacc, 0, (size_t)g.nx * g.nz * sizeof(double), S()),
                           "memset");
  return rc;
}

// ------------------------------------------------- binomial checkpointing
//
// The backward pass needs the state before each step, in reverse order.
// With s snapshot slots beyond the initial state, reversing n steps costs
// cost(n, s) forward steps, the minimum over the first snapshot position m
// of m + cost(n - m, s - 1) + cost(m, s), with cost(1, s) = 0 and
// cost(n, 0) = n (n - 1) / 2 (every state recomputed from the range's
// start).  The recursion is the binomial checkpointing of Griewank and
// Walther (ACM TOMS 26, 2000): its optimum reverses C(s + r, s) steps with
// r forward sweeps.  Solved exactly by dynamic programming on the host.
struct Schedule {
  int n = 0, s = 0;
  std::vector<long long> cost;
  std::vector<int> split;
  long long& C(int len, int sl) { return cost[(size_t)len * (s + 1) + sl]; }
  int& M(int len, int sl) { return split[(size_t)len * (s + 1) + sl]; }
  void build(int n_, int s_) {
    n = n_;
    s = s_;
    cost.assign((size_t)(n + 1) * (s + 1), 0);
    split.assign((size_t)(n + 1) * (s + 1), 0);
    for (int len = 1; len <= n; len++) {
      C(len, 0) = (long long)len * (len - 1) / 2;
      for (int sl = 1; sl <= s; sl++) {
        if (len == 1) {
          C(len, sl) = 0;
          continue;
        }
        long long best = -1;
        int bm = 1;
        for (int m = 1; m < len; m++) {
          const long long c = m + C(len - m, sl - 1) + C(m, sl);
          if (best < 0 || c < best) {
            best = c;
            bm = m;
          }
        }
        C(len, sl) = best;
        M(len, sl) = bm;
      }
    }
  }
};

// The forward trajectory the backward pass revisits.  FP64: the state of
// vk_time_step.  Mixed: the state of vk_time_step_mixed -- u, v, w (and
// p, u_temp.., divu) in FP64; T, the AB2 history and u_ta in float, with
// float copies of cellvalue, rhoCp and thc -- about which the backward
// pass linearises the FP64 step.  Snapshots hold the 11 state fields
// (mixed: 3 double and 8 float); the initial state is read from the
// caller's arrays, so it takes no slot.
enum { MT, MUP, MVP, MWP, MTP, MUTA, MVTA, MWTA, kMixedF };

struct Traj {
  const Problem* P = nullptr;
  const Inputs* in = nullptr;
  bool mixed = false;
  Work* w = nullptr;  // the FP64 state (the trajectory's own when FP64)
  // the mixed trajectory's state
  double *mu = nullptr, *mv = nullptr, *mw = nullptr, *mp = nullptr, *mut = nullptr,
         *mvt = nullptr, *mwt = nullptr, *mdiv = nullptr;
  float* mf[kMixedF] = {};
  float *cvf = nullptr, *rcf = nullptr, *thcf = nullptr;
  std::vector<void*> snap;  // slot q: kState consecutive pointers
  long long steps_run = 0;

  size_t nc() const { return P->g.cells(); }
  void copyd(const double* a, double* b) const { copy(P->g, a, b); }
  void copyf(const float* a, float* b) const {
    const long long n = (long long)nc();
    k_copy_cells_f<<<(unsigned)((n + 255) / 256), 256, 0, S()>>>(n, a, b);
  }
  void tod(const float* a, double* b, long long n) const {
    k_to_double<<<(unsigned)((n + 255) / 256), 256, 0, S()>>>(n, a, b);
  }
  void tof(const double* a, float* b, long long n) const {
    k_to_float<<<(unsigned)((n + 255) / 256), 256, 0, S()>>>(n, a, b);
  }
  // bytes of one snapshot
  size_t snap_bytes() const {
    return mixed ? nc() * (3 * sizeof(double) + kMixedF * sizeof(float))
                 : nc() * kState * sizeof(double);
  }
  int alloc(Dev& dev, int slots) {
    if (mixed) {
      double** d[8] = {&mu, &mv, &mw, &mp, &mut, &mvt, &mwt, &mdiv};
      for (auto q : d)
        if (!(*q = dev.alloc(nc()))) return fail(VK_ERR_HIP, "adjoint: out of device memory");
      for (auto& q : mf)
        if (!(q = dev.allocf(nc()))) return fail(VK_ERR_HIP, "adjoint: out of device memory");
      cvf = dev.allocf(P->g.interior() * 4);
      rcf = dev.allocf(nc());
      thcf = dev.allocf(nc());
      if (!cvf || !rcf || !thcf) return fail(VK_ERR_HIP, "adjoint: out of device memory");
    }
    snap.assign((size_t)slots * kState, nullptr);
    for (int q = 0; q < slots; q++)
      for (int f = 0; f < kState; f++) {
        const bool dbl = !mixed || f < 3;
        void* p = dbl ? (void*)dev.alloc(nc()) : (void*)dev.allocf(nc());
        if (!p) return fail(VK_ERR_HIP, "adjoint: out of device memory for the snapshots");
        snap[(size_t)q * kState + f] = p;
      }
    return VK_OK;
  }
  // the trajectory state := the caller's initial state
  void load_initial() {
    const double* src[kState] = {in->u,  in->v,  in->w,   in->t,   in->up, in->vp,
                                 in->wp, in->tp, in->uta, in->vta, in->wta};
    if (!mixed) {
      for (int f = 0; f < kState; f++) copyd(src[f], w->st[f]);
      copyd(in->p, w->p);
      return;
    }
    copyd(in->u, mu);
    copyd(in->v, mv);
    copyd(in->w, mw);
    copyd(in->p, mp);
    for (int f = 3; f < kState; f++) tof(src[f], mf[f - 3], (long long)nc());
  }
  // state arrays of the trajectory, field f of kState
  void* field(int f) {
    if (!mixed) return w->st[f];
    if (f < 3) return f == 0 ? (void*)mu : f == 1 ? (void*)mv : (void*)mw;
    return mf[f - 3];
  }
  void save(int slot) {
    for (int f = 0; f < kState; f++) {
      void* d = snap[(size_t)slot * kState + f];
      if (!mixed || f < 3)
        copyd(static_cast<double*>(field(f)), static_cast<double*>(d));
      else
        copyf(static_cast<float*>(field(f)), static_cast<float*>(d));
    }
  }
  void load(int slot) {
    if (slot < 0) return load_initial();
    for (int f = 0; f < kState; f++) {
      const void* d = snap[(size_t)slot * kState + f];
      if (!mixed || f < 3)
        copyd(static_cast<const double*>(d), static_cast<double*>(field(f)));
      else
        copyf(static_cast<const float*>(d), static_cast<float*>(field(f)));
    }
  }
  // one step of the trajectory (state before step l -> after it)
  int advance(int l) {
    steps_run++;
    if (!mixed) return forward_step(*P, *w, l);
    const Problem& Q = *P;
    const Geom& g = Q.g;
    return vk_time_step_mixed(mu, mv, mw, mp, mf[MT], mf[MUP], mf[MVP], mf[MWP], mf[MTP],
                              mf[MUTA], mf[MVTA], mf[MWTA], mut, mvt, mwt, mdiv, cvf, rcf, thcf,
                              Q.m.dx, Q.m.dy, Q.m.dz, Q.m.dx12, Q.m.dy12, Q.m.dz12, &g.nx, &g.ny,
                              &g.nz, &Q.dt, &Q.re, &Q.sc, &Q.eta, &Q.xl, &Q.zl, &l);
  }
  void objective_acc() {
    const int n = P->g.nx * P->g.nz;
    if (!mixed)
      k_objective_acc<double><<<(n + 255) / 256, 256, 0, S()>>>(P->g, P->m, w->st[ST], P->dt,
                                                                w->jacc);
    else
      k_objective_acc<float><<<(n + 255) / 256, 256, 0, S()>>>(P->g, P->m, mf[MT], P->dt,
                                                               w->jacc);
  }
  // the FP64 work state := the trajectory state (the linearisation point
  // of the taped step); the input u_ta kept in uta[3]
  void to_work(double* const* uta) {
    if (mixed) {
      copyd(mu, w->st[SU]);
      copyd(mv, w->st[SV]);
      copyd(mw, w->st[SW]);
      copyd(mp, w->p);
      for (int f = 3; f < kState; f++) tod(mf[f - 3], w->st[f], (long long)nc());
    }
    for (int c = 0; c < 3; c++) copyd(w->st[SUTA + c], uta[c]);
  }
};

// Runs the schedule: reverses steps a+1..a+n given the state after step a
// in `slot` (-1: the initial state) and the free slots from `next` on.
// `adjoint(l)` runs the backward step l from the trajectory state before it.
template <typename ADJ>
int reverse(Traj& T, Schedule& sc, int a, int n, int slot, int next, int free_slots, ADJ& adjoint) {
  int rc;
  if (n == 1) {
    T.load(slot);
    return adjoint(a + 1);
  }
  if (free_slots == 0) {
    for (int i = n - 1; i >= 0; i--) {
      T.load(slot);
      for (int l = a + 1; l <= a + i; l++)
        if ((rc = T.advance(l))) return rc;
      if ((rc = adjoint(a + i + 1))) return rc;
    }
    return VK_OK;
  }
  const int m = sc.M(n, free_slots);
  T.load(slot);
  for (int l = a + 1; l <= a + m; l++)
    if ((rc = T.advance(l))) return rc;
  T.save(next);
  if ((rc = reverse(T, sc, a + m, n - m, next, next + 1, free_slots - 1, adjoint))) return rc;
  return reverse(T, sc, a, m, slot, next, free_slots, adjoint);
}

}  // namespace

extern "C" {

int vk_objective(const double* u, const double* v, const double* w, const double* p,
                 const double* t, const double* u_pre, const double* v_pre, const double* w_pre,
                 const double* t_pre, const double* u_ta, const double* v_ta, const double* w_ta,
                 const double* cellvalue, const double* rhocp, const double* thc,
                 const double* dx, const double* dy, const double* dz, const double* dx12,
                 const double* dy12, const double* dz12, const int* nx, const int* ny,
                 const int* nz, const double* dt, const double* re, const double* sc,
                 const double* eta, const double* xl, const double* zl, const int* nsteps,
                 double* objective) {
  VK_REQUIRE_INIT();
  if (*nsteps < 1) return fail(VK_ERR_ARGUMENT, "objective: nsteps >= 1");
  Problem P{Geom(*nx, *ny, *nz), Metrics{}, *dt, *re, *sc, *eta, *xl, *zl, nullptr, nullptr,
            nullptr};
  if (P.g.nx < 4 || P.g.ny < 4 || P.g.nz < 4) return fail(VK_ERR_BAD_SIZE, "objective: n >= 4");
  if (!P.g.fits32()) return fail(VK_ERR_BAD_SIZE, "objective: grid too large");
  if (int frc = fft_sizes_ok(P.g)) return frc;
  Inputs in;
  int rc = resolve_problem(P, in, u, v, w, p, t, u_pre, v_pre, w_pre, t_pre, u_ta, v_ta, w_ta,
                           cellvalue, rhocp, thc, dx, dy, dz, dx12, dy12, dz12);
  if (rc) return rc;
  Dev dev;
  P.cv = dev.alloc(P.g.interior() * 4);
  if (!P.cv) return fail(VK_ERR_HIP, "objective: out of device memory");
  rc = hip_check(hipMemcpyAsync(P.cv, in.cv, P.g.interior() * 4 * sizeof(double),
                                hipMemcpyDeviceToDevice, S()),
                 "cellvalue copy");
  if (rc) return rc;
  Work wk;
  rc = init_work(P, in, dev, wk);
  if (rc) return rc;
  for (int l = 1; l <= *nsteps; l++) {
    rc = forward_step(P, wk, l);
    if (rc) return rc;
    objective_acc(P, wk);
  }
  return sum_objective(P.g, wk.jacc, objective);
}

int vk_adjoint_gradient(const double* u, const double* v, const double* w, const double* p,
                        const double* t, const double* u_pre, const double* v_pre,
                        const double* w_pre, const double* t_pre, const double* u_ta,
                        const double* v_ta, const double* w_ta, const double* cellvalue,
                        const double* rhocp, const double* thc, const double* dx,
                        const double* dy, const double* dz, const double* dx12,
                        const double* dy12, const double* dz12, const int* nx, const int* ny,
                        const int* nz, const double* dt, const double* re, const double* sc,
                        const double* eta, const double* xl, const double* zl,
                        const int* nsteps, const int* checkpoint_every, double* objective,
                        double* grad) {
  VK_REQUIRE_INIT();
  const int N = *nsteps, C = *checkpoint_every;
  if (N < 1 || C < 1) return fail(VK_ERR_ARGUMENT, "adjoint: nsteps, checkpoint_every >= 1");
  Problem P{Geom(*nx, *ny, *nz), Metrics{}, *dt, *re, *sc, *eta, *xl, *zl, nullptr, nullptr,
            nullptr};
  const Geom& g = P.g;
  if (g.nx < 4 || g.ny < 4 || g.nz < 4) return fail(VK_ERR_BAD_SIZE, "adjoint: n >= 4");
  if (!g.fits32()) return fail(VK_ERR_BAD_SIZE, "adjoint: grid too large");
  if (int frc = fft_sizes_ok(g)) return frc;
  Inputs in;
  int rc = resolve_problem(P, in, u, v, w, p, t, u_pre, v_pre, w_pre, t_pre, u_ta, v_ta, w_ta,
                           cellvalue, rhocp, thc, dx, dy, dz, dx12, dy12, dz12);
  if (rc) return rc;
  double* dgrad = static_cast<double*>(resolve(grad, g.interior() * sizeof(double), "grad"));
  if (!dgrad) return last_status();
  Dev dev;
  const size_t nc = g.cells();
  P.cv = dev.alloc(g.interior() * 4);
  if (!P.cv) return fail(VK_ERR_HIP, "adjoint: out of device memory");
  rc = hip_check(hipMemcpyAsync(P.cv, in.cv, g.interior() * 4 * sizeof(double),
                                hipMemcpyDeviceToDevice, S()),
                 "cellvalue copy");
  if (rc) return rc;
  Work wk;
  rc = init_work(P, in, dev, wk);
  if (rc) return rc;

  // forward with checkpoints: cp[s] = state before step s*C + 1
  const int ncp = (N + C - 1) / C;
  std::vector<double*> cp((size_t)ncp * kState);
  for (auto& q : cp)
    if (!(q = dev.alloc(nc))) return fail(VK_ERR_HIP, "adjoint: out of device memory");
  for (int l = 1; l <= N; l++) {
    if ((l - 1) % C == 0)
      for (int f = 0; f < kState; f++) copy(g, wk.st[f], cp[(size_t)((l - 1) / C) * kState + f]);
    rc = forward_step(P, wk, l);
    if (rc) return rc;
    objective_acc(P, wk);
  }
  rc = sum_objective(g, wk.jacc, objective);
  if (rc) return rc;

  // backward buffers.  Where a segment's tapes fit in device memory with
  // 4 GiB to spare (per step the 16 tape fields, the 4 inputs, the 3 u_ta
  // and the dirty-line masks), every step of a segment is taped once, as
  // the segment is recomputed from its checkpoint, and the backward pass
  // runs on those tapes.  Otherwise the recompute keeps each step's state
  // and the backward pass reruns each step with one step's tape.
  const size_t nxl = (size_t)g.ny * g.nz, nyl = (size_t)g.nx * g.nz;
  const size_t nflags = nxl + nyl + (size_t)g.nx * g.ny;
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
    (void)hipGetLastError();
    free_b = 0;
  }
  const size_t keep_bytes = (size_t)C * (23 * nc * sizeof(double) + nflags * sizeof(unsigned));
  // VK_ADJOINT_SEGMENT_TAPES=0 in the environment asks for the per-step
  // path whatever the memory (the two give the same gradient bit for bit)
  const char* seg_env = std::getenv("VK_ADJOINT_SEGMENT_TAPES");
  const bool keep_tapes = keep_bytes + ((size_t)4 << 30) < free_b &&
                          !(seg_env && std::strcmp(seg_env, "0") == 0);
  std::vector<double*> seg, tseg, iseg, useg;
  unsigned* fseg = nullptr;
  double* tape[16] = {};
  double* in4[4] = {};
  if (keep_tapes) {
    tseg.resize((size_t)C * 16);
    iseg.resize((size_t)C * 4);
    useg.resize((size_t)C * 3);
    for (auto* v : {&tseg, &iseg, &useg})
      for (auto& q : *v)
        if (!(q = dev.alloc(nc))) return fail(VK_ERR_HIP, "adjoint: out of device memory");
    fseg = static_cast<unsigned*>(dev.raw((size_t)C * nflags * sizeof(unsigned)));
    if (!fseg) return fail(VK_ERR_HIP, "adjoint: out of device memory");
  } else {
    seg.resize((size_t)C * kState);
    for (auto& q : seg)
      if (!(q = dev.alloc(nc))) return fail(VK_ERR_HIP, "adjoint: out of device memory");
    for (auto& q : tape)
      if (!(q = dev.alloc(nc))) return fail(VK_ERR_HIP, "adjoint: out of device memory");
    for (auto& q : in4)
      if (!(q = dev.alloc(nc))) return fail(VK_ERR_HIP, "adjoint: out of device memory");
  }
  AdjBufs B;
  for (int f = 0; f < kState; f++) {
    B.lam[f] = dev.alloc(nc);
    B.nxt[f] = dev.alloc(nc);
    if (!B.lam[f] || !B.nxt[f]) return fail(VK_ERR_HIP, "adjoint: out of device memory");
    rc = hip_check(hipMemsetAsync(B.lam[f], 0, nc * sizeof(double), S()), "memset");
    if (rc) return rc;
  }
  for (int c = 0; c < 3; c++) {
    B.lut[c] = dev.alloc(nc);
    B.lg[c] = dev.alloc(nc);
  }
  B.lphi = dev.alloc(nc);
  B.ysol = dev.alloc(nc);
  B.spec = dev.alloc(2 * g.spec_cells());
  bool w8ok = true;
  for (auto& q : B.w8) w8ok = (q = dev.alloc(nc)) != nullptr && w8ok;
  B.coef = dev.alloc((size_t)(g.nx / 2 + 1) * (g.nz + 1) + 3 * (size_t)g.ny + 8);
  B.gcv = dev.alloc(g.interior() * 4);
  if (!B.lut[0] || !B.lut[1] || !B.lut[2] || !B.lg[0] || !B.lg[1] || !B.lg[2] || !B.lphi ||
      !B.ysol || !B.spec || !w8ok || !B.coef || !B.gcv)
    return fail(VK_ERR_HIP, "adjoint: out of device memory");
  rc = hip_check(hipMemsetAsync(B.gcv, 0, g.interior() * 4 * sizeof(double), S()), "memset");
  if (rc) return rc;
  const int nJ = g.nx * g.nz;
  // the adjoint of the final state: the objective's dependence on T^N
  k_objective_src<<<(nJ + 255) / 256, 256, 0, S()>>>(g, P.m, P.dt, B.lam[ST]);

  for (int sidx = ncp - 1; sidx >= 0; sidx--) {
    const int l0 = sidx * C + 1, l1 = std::min(N, (sidx + 1) * C);
    for (int f = 0; f < kState; f++) copy(g, cp[(size_t)sidx * kState + f], wk.st[f]);
    if (keep_tapes) {
      // every step of the segment taped from its checkpoint, with its
      // input u_ta and a copy of its dirty-line masks
      for (int l = l0; l <= l1; l++) {
        const size_t q = (size_t)(l - l0);
        for (int c = 0; c < 3; c++) copy(g, wk.st[SUTA + c], useg[q * 3 + c]);
        LineFlags lf{};
        rc = taped_step(P, wk, l, tseg.data() + q * 16, iseg.data() + q * 4, &lf);
        if (rc) return rc;
        rc = hip_check(hipMemcpyAsync(fseg + q * nflags, lf.fx, nflags * sizeof(unsigned),
                                      hipMemcpyDeviceToDevice, S()),
                       "adjoint masks");
        if (rc) return rc;
      }
      for (int l = l1; l >= l0; l--) {
        const size_t q = (size_t)(l - l0);
        const double* uta_in[3] = {useg[q * 3], useg[q * 3 + 1], useg[q * 3 + 2]};
        unsigned* fq = fseg + q * nflags;
        const LineFlags lf{fq, fq + nxl, fq + nxl + nyl};
        rc = adjoint_step(P, B, l, tseg.data() + q * 16, iseg.data() + q * 4, uta_in, &lf);
        if (rc) return rc;
        if (l > 1) k_objective_src<<<(nJ + 255) / 256, 256, 0, S()>>>(g, P.m, P.dt, B.lam[ST]);
      }
      continue;
    }
    // recompute the segment from its checkpoint, keeping each step's input
    for (int l = l0; l <= l1; l++) {
      for (int f = 0; f < kState; f++) copy(g, wk.st[f], seg[(size_t)(l - l0) * kState + f]);
      if (l < l1) {
        rc = forward_step(P, wk, l);
        if (rc) return rc;
      }
    }
    // backward through the segment
    for (int l = l1; l >= l0; l--) {
      for (int f = 0; f < kState; f++) copy(g, seg[(size_t)(l - l0) * kState + f], wk.st[f]);
      const double* uta_in[3] = {seg[(size_t)(l - l0) * kState + SUTA],
                                 seg[(size_t)(l - l0) * kState + SVTA],
                                 seg[(size_t)(l - l0) * kState + SWTA]};
      LineFlags lf{};
      rc = taped_step(P, wk, l, tape, in4, &lf);
      if (rc) return rc;
      rc = adjoint_step(P, B, l, tape, in4, uta_in, &lf);
      if (rc) return rc;
      if (l > 1) k_objective_src<<<(nJ + 255) / 256, 256, 0, S()>>>(g, P.m, P.dt, B.lam[ST]);
    }
  }
  k_adj_cv_faces<<<interior_grid(g), interior_block(), 0, S()>>>(g, B.gcv, dgrad);
  rc = hip_check(hipStreamSynchronize(S()), "adjoint");
  if (rc) return rc;
  VK_LAUNCH_CHECK("adjoint_gradient");
  return VK_OK;
}

int vk_adjoint_gradient_revolve(
    const double* u, const double* v, const double* w, const double* p, const double* t,
    const double* u_pre, const double* v_pre, const double* w_pre, const double* t_pre,
    const double* u_ta, const double* v_ta, const double* w_ta, const double* cellvalue,
    const double* rhocp, const double* thc, const double* dx, const double* dy, const double* dz,
    const double* dx12, const double* dy12, const double* dz12, const int* nx, const int* ny,
    const int* nz, const double* dt, const double* re, const double* sc, const double* eta,
    const double* xl, const double* zl, const int* nsteps, const int* snapshots,
    const int* mixed, double* objective, double* grad, double* stats) {
  VK_REQUIRE_INIT();
  const int N = *nsteps;
  if (N < 1) return fail(VK_ERR_ARGUMENT, "adjoint: nsteps >= 1");
  Problem P{Geom(*nx, *ny, *nz), Metrics{}, *dt, *re, *sc, *eta, *xl, *zl, nullptr, nullptr,
            nullptr};
  const Geom& g = P.g;
  if (g.nx < 4 || g.ny < 4 || g.nz < 4) return fail(VK_ERR_BAD_SIZE, "adjoint: n >= 4");
  if (!g.fits32()) return fail(VK_ERR_BAD_SIZE, "adjoint: grid too large");
  if (int frc = fft_sizes_ok(g)) return frc;
  Inputs in;
  int rc = resolve_problem(P, in, u, v, w, p, t, u_pre, v_pre, w_pre, t_pre, u_ta, v_ta, w_ta,
                           cellvalue, rhocp, thc, dx, dy, dz, dx12, dy12, dz12);
  if (rc) return rc;
  double* dgrad = static_cast<double*>(resolve(grad, g.interior() * sizeof(double), "grad"));
  if (!dgrad) return last_status();
  Dev dev;
  const size_t nc = g.cells();
  P.cv = dev.alloc(g.interior() * 4);
  if (!P.cv) return fail(VK_ERR_HIP, "adjoint: out of device memory");
  rc = hip_check(hipMemcpyAsync(P.cv, in.cv, g.interior() * 4 * sizeof(double),
                                hipMemcpyDeviceToDevice, S()),
                 "cellvalue copy");
  if (rc) return rc;
  Work wk;
  rc = init_work(P, in, dev, wk);
  if (rc) return rc;
  // the backward step's buffers
  double* tape[16];
  double* in4[4];
  double* uta[3];
  for (auto& q : tape)
    if (!(q = dev.alloc(nc))) return fail(VK_ERR_HIP, "adjoint: out of device memory");
  for (auto& q : in4)
    if (!(q = dev.alloc(nc))) return fail(VK_ERR_HIP, "adjoint: out of device memory");
  for (auto& q : uta)
    if (!(q = dev.alloc(nc))) return fail(VK_ERR_HIP, "adjoint: out of device memory");
  AdjBufs B;
  for (int f = 0; f < kState; f++) {
    B.lam[f] = dev.alloc(nc);
    B.nxt[f] = dev.alloc(nc);
    if (!B.lam[f] || !B.nxt[f]) return fail(VK_ERR_HIP, "adjoint: out of device memory");
  }
  for (int c = 0; c < 3; c++) {
    B.lut[c] = dev.alloc(nc);
    B.lg[c] = dev.alloc(nc);
  }
  B.lphi = dev.alloc(nc);
  B.ysol = dev.alloc(nc);
  B.spec = dev.alloc(2 * g.spec_cells());
  bool w8ok = true;
  for (auto& q : B.w8) w8ok = (q = dev.alloc(nc)) != nullptr && w8ok;
  B.coef = dev.alloc((size_t)(g.nx / 2 + 1) * (g.nz + 1) + 3 * (size_t)g.ny + 8);
  B.gcv = dev.alloc(g.interior() * 4);
  if (!B.lut[0] || !B.lut[1] || !B.lut[2] || !B.lg[0] || !B.lg[1] || !B.lg[2] || !B.lphi ||
      !B.ysol || !B.spec || !w8ok || !B.coef || !B.gcv)
    return fail(VK_ERR_HIP, "adjoint: out of device memory");
  Traj T;
  T.P = &P;
  T.in = &in;
  T.mixed = *mixed != 0;
  T.w = &wk;
  rc = T.alloc(dev, 0);
  if (rc) return rc;
  if (T.mixed) {
    T.tof(P.cv, T.cvf, (long long)g.interior() * 4);
    T.tof(in.rc, T.rcf, (long long)nc);
    T.tof(in.thc, T.thcf, (long long)nc);
  }
  const double* uta_in[3] = {uta[0], uta[1], uta[2]};
  const int nJ = g.nx * g.nz;
  auto zero = [&](double* a, size_t n) {
    return hip_check(hipMemsetAsync(a, 0, n * sizeof(double), S()), "memset");
  };
  // with VK_ADJOINT_PROGRESS set in the environment, a line on stderr
  // every twentieth of the backward pass (after the work so far has run)
  const bool progress = std::getenv("VK_ADJOINT_PROGRESS") != nullptr;
  const int every = std::max(1, N / 20);
  int done = 0;
  auto report = [&](const char* what, int k) -> int {
    int r = hip_check(hipStreamSynchronize(S()), "adjoint progress");
    if (r) return r;
    std::fprintf(stderr, "[adjoint] %s %d of %d (forward steps run %lld)\n", what, k, N,
                 T.steps_run);
    std::fflush(stderr);
    return VK_OK;
  };
  auto adjoint = [&](int l) -> int {
    T.to_work(uta);
    LineFlags lf{};
    int r = taped_step(P, wk, l, tape, in4, &lf);
    if (r) return r;
    r = adjoint_step(P, B, l, tape, in4, uta_in, &lf);
    if (r) return r;
    if (l > 1) k_objective_src<<<(nJ + 255) / 256, 256, 0, S()>>>(g, P.m, P.dt, B.lam[ST]);
    done++;
    if (progress && done % every == 0) return report("backward step", done);
    return VK_OK;
  };
  // one step of each kind first, so that every workspace the steps grow
  // exists before the snapshots take what memory is left
  T.load_initial();
  if ((rc = T.advance(1))) return rc;
  T.load_initial();
  for (int f = 0; f < kState; f++)
    if ((rc = zero(B.lam[f], nc))) return rc;
  if ((rc = zero(B.gcv, g.interior() * 4))) return rc;
  if ((rc = adjoint(1))) return rc;
  T.steps_run = 0;
  // the snapshots: as asked, or as many as the free memory holds (keeping
  // 4 GiB), never more than the schedule can use
  int s = *snapshots;
  if (s < 0) {
    size_t fr = 0, tot = 0;
    if ((rc = hip_check(hipMemGetInfo(&fr, &tot), "hipMemGetInfo"))) return rc;
    const size_t keep = (size_t)4 << 30;
    s = fr > keep ? (int)std::min<size_t>((fr - keep) / T.snap_bytes(), 1u << 20) : 0;
  }
  s = std::min(s, std::max(N - 1, 0));
  {  // the snapshot slots
    T.snap.assign((size_t)s * kState, nullptr);
    for (int q = 0; q < s; q++)
      for (int f = 0; f < kState; f++) {
        const bool dbl = !T.mixed || f < 3;
        void* pp = dbl ? (void*)dev.alloc(nc) : (void*)dev.allocf(nc);
        if (!pp)
          return fail(VK_ERR_HIP, "adjoint: out of device memory for " + std::to_string(s) +
                                      " snapshots");
        T.snap[(size_t)q * kState + f] = pp;
      }
  }
  Schedule sch;
  sch.build(N, s);
  // the objective over the whole trajectory
  T.load_initial();
  if ((rc = zero(wk.jacc, (size_t)nJ))) return rc;
  for (int l = 1; l <= N; l++) {
    if ((rc = T.advance(l))) return rc;
    T.objective_acc();
    if (progress && l % every == 0 && (rc = report("objective step", l))) return rc;
  }
  if ((rc = sum_objective(g, wk.jacc, objective))) return rc;
  // backward
  for (int f = 0; f < kState; f++)
    if ((rc = zero(B.lam[f], nc))) return rc;
  if ((rc = zero(B.gcv, g.interior() * 4))) return rc;
  k_objective_src<<<(nJ + 255) / 256, 256, 0, S()>>>(g, P.m, P.dt, B.lam[ST]);
  if ((rc = reverse(T, sch, 0, N, -1, 0, s, adjoint))) return rc;
  k_adj_cv_faces<<<interior_grid(g), interior_block(), 0, S()>>>(g, B.gcv, dgrad);
  rc = hip_check(hipStreamSynchronize(S()), "adjoint");
  if (rc) return rc;
  VK_LAUNCH_CHECK("adjoint_gradient_revolve");
  if (stats) {
    stats[0] = s;
    stats[1] = (double)T.steps_run;
    stats[2] = (double)sch.C(N, s);
    stats[3] = (double)dev.bytes;
  }
  return VK_OK;
}

}  // extern "C"
