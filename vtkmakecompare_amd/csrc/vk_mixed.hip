// vk_mixed.hip -- the mixed-precision form of the fused time step.
//
// Storage: the velocity state u, v, w, the ADI results u_temp, v_temp,
// w_temp, the pressure p and divu are FP64; temperature, the AB2 history,
// u_ta, cellvalue, rhoCp and thc are float.  Arithmetic:
//   - the explicit phase (convection, diffusion, penalisation, AB2 and the
//     temperature update) runs in FP32 with FP32 accumulation; the
//     velocity increment is added to the FP64 velocity and divided by
//     (1 + eta cv dt) in FP64, so the state never passes through float;
//   - the ADI line solves eliminate in FP64 on their mixed storage;
//   - the projection runs float FFTs and spectra, FP64 tridiagonal
//     eliminations (SPN, SP0B), and one iterative-refinement pass whose
//     residual is formed in FP64 (vk_mkp.hip: launch_mkp_mixed).
// This unit is built with FP contraction on: nothing here has to
// reproduce the reference's rounding bit for bit.
#include "vk_kernels.h"
#include "vk_runtime.h"
#include "vk_stencil_ops.h"

namespace vk {

// float copies of the metric arrays and their reciprocals; one struct
// serves as both metric sets of the stencil templates
struct FMet {
  const float *dx, *dy, *dz, *rdx, *rdy, *rdz, *rdx12, *rdy12, *rdz12;
};

__global__ void k_fmetrics(Metrics m, int nx, int ny, int nz, float* out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int ax = nx + 2, ay = ny + 2, az = nz + 2;
  float* p = out;
  if (t < ax) {
    p[t] = (float)m.dx[t];
    p[3 * ax + t] = (float)(1.0 / m.dx[t]);
    p[6 * ax + t] = (float)(1.0 / m.dx12[t]);
  }
  p += 9 * ax;
  if (t < ay) {
    p[t] = (float)m.dy[t];
    p[3 * ay + t] = (float)(1.0 / m.dy[t]);
    p[6 * ay + t] = (float)(1.0 / m.dy12[t]);
  }
  p += 9 * ay;
  if (t < az) {
    p[t] = (float)m.dz[t];
    p[3 * az + t] = (float)(1.0 / m.dz[t]);
    p[6 * az + t] = (float)(1.0 / m.dz12[t]);
  }
}

// Layout of the float metric block: per axis a of length n_a = n + 2, nine
// slots of which 0, 3 and 6 hold d, 1/d and 1/d12.
static int make_fmetrics(const Geom& g, const Metrics& m, FMet& f, hipStream_t s) {
  const int ax = g.nx + 2, ay = g.ny + 2, az = g.nz + 2;
  float* w = static_cast<float*>(workspace(kWorkRecipF, sizeof(float) * 9 * (ax + ay + az)));
  if (!w) return VK_ERR_HIP;
  int nmax = ax > ay ? ax : ay;
  nmax = nmax > az ? nmax : az;
  k_fmetrics<<<(nmax + 255) / 256, 256, 0, s>>>(m, g.nx, g.ny, g.nz, w);
  const float* x = w;
  const float* y = w + 9 * ax;
  const float* z = y + 9 * ay;
  f.dx = x;
  f.rdx = x + 3 * ax;
  f.rdx12 = x + 6 * ax;
  f.dy = y;
  f.rdy = y + 3 * ay;
  f.rdy12 = y + 6 * ay;
  f.dz = z;
  f.rdz = z + 3 * az;
  f.rdz12 = z + 6 * az;
  return VK_OK;
}

struct ExplicitMixedArgs {
  const double *u, *v, *w;
  const float *t, *rc, *thc, *u_ta, *v_ta, *w_ta, *cv;
  float *u_pre, *v_pre, *w_pre, *t_pre;
  double *ua, *va, *wa;
  float* ut;
  float dt, eta, a1, a2, b1, c0, cd;
  double dtd, etad;
  LineFlags lf;  // filled when non-null
};

// x / (1 + eta cv dt): exactly x where cv is 0 (the divisor is exactly 1)
template <typename R>
__device__ inline R pen_div_m(R x, R eta, R cv, R dt) {
  return x / (R(1) + eta * cv * dt);
}

// The explicit phase of one step: the same terms as k_explicit_fused
// (vk_step.hip), from the FP64 velocity read as float.
// u, v, w, then T, in one pass: T takes the velocity at the point and its
// lower neighbours from the stencils loaded for u, v, w, and T's loads are
// issued before the velocity's stores
__global__ __launch_bounds__(256) void k_explicit_mixed(Geom g, FMet f, ExplicitMixedArgs x) {
  int bx, by, bz;
  xcd_block(bx, by, bz);
  const int i = 1 + bx * blockDim.x + threadIdx.x;
  const int j = 1 + by * blockDim.y + threadIdx.y;
  const int k = 1 + bz;
  if (i > g.nx || j > g.ny) return;
  const long long o = g.at(i, j, k);
  const float dt = x.dt, eta = x.eta;
  {
    const float cv2 = x.cv[g.cv(i, j, k, 2)], cv3 = x.cv[g.cv(i, j, k, 3)],
                cv4 = x.cv[g.cv(i, j, k, 4)];
    const float cv1 = x.cv[g.cv(i, j, k, 1)];
    // the three stencils and the history, every load issued before any
    // arithmetic; the schedule barrier keeps the compiler from spreading
    // them over the computation
    Nb27<double, float> U, V, W;
    load_velocity_nbhd(U, V, W, x.u, x.v, x.w, g, i, j, k);
    const float upre = x.u_pre[o], vpre = x.v_pre[o], wpre = x.w_pre[o];
    PtMet<float> M;
    M.load_f(f, i, j, k);
    __builtin_amdgcn_sched_barrier(0);
    unsigned mk = x.lf.fx ? dirty_bits(0.0f, cv2, cv3, cv4, 1.0f, 1.0f) : 0u;
    // u_ta read only where its term eta cv u_ta dt is not 0
    float uta = 0.0f, vta = 0.0f, wta = 0.0f;
    if (cv2 != 0.0f) uta = x.u_ta[o];
    if (cv3 != 0.0f) vta = x.v_ta[o];
    if (cv4 != 0.0f) wta = x.w_ta[o];
    // every value computed before the first store (the stores would fence
    // the later loads: the outputs may alias the inputs as far as the
    // compiler knows)
    const float uv = conv_u_r<float>(U, V, W, M, M, i, j, k, dt);
    const float ud = cond_u_r<float>(U, M, i, j, k, x.c0);
    const float ul = eta * cv2 * uta * dt;
    const float ua = x.a1 * uv + x.a2 * upre + x.b1 * ud + x.b1 * ul;
    const float vv = conv_v_r<float>(U, V, W, M, M, i, j, k, dt);
    const float vd = cond_v_r<float>(V, M, i, j, k, x.c0);
    const float vl = eta * cv3 * vta * dt;
    const float va = x.a1 * vv + x.a2 * vpre + x.b1 * vd + x.b1 * vl;
    const float wv = conv_w_r<float>(U, V, W, M, M, i, j, k, dt);
    const float wd = cond_w_r<float>(W, M, i, j, k, x.c0);
    const float wl = eta * cv4 * wta * dt;
    const float wa = x.a1 * wv + x.a2 * wpre + x.b1 * wd + x.b1 * wl;
    const double un = pen_div_m<double>(U.v[1][1][1] + (double)ua, x.etad, (double)cv2, x.dtd);
    const double vn = pen_div_m<double>(V.v[1][1][1] + (double)va, x.etad, (double)cv3, x.dtd);
    const double wn = pen_div_m<double>(W.v[1][1][1] + (double)wa, x.etad, (double)cv4, x.dtd);
    // T's 7-point stencils of T, rhoCp and thc
    Nb27<float, float> T, RC, K;
    load_scalar_fields(T, RC, K, x.t, x.rc, x.thc, g, i, j, k);
    const float tpre = x.t_pre[o];
    x.u_pre[o] = uv;
    x.ua[o] = un;
    x.v_pre[o] = vv;
    x.va[o] = vn;
    // SET_EXPLICIT's reflection v_a(i,ny+1,k) = -v_a(i,ny,k) of the
    // increment (nse_dominik.f:1132-1286)
    if (j == g.ny) x.va[g.at(i, g.ny + 1, k)] = -(double)va;
    x.w_pre[o] = wv;
    x.wa[o] = wn;
    // T, with the velocity values of U, V, W
    __builtin_amdgcn_sched_barrier(0);
    const float rc0 = RC(i, j, k);
    if (x.lf.fx) mk |= dirty_bits(cv1, 0.0f, 0.0f, 0.0f, rc0, K(i, j, k));
    const float rrc = 1.0f / rc0;
    const float vt = conv_scalar_r<float>(U, V, W, T, RC, M, M, i, j, k, dt, rrc);
    const float dtt = diff_scalar_r<float>(T, K, M, M, i, j, k, x.cd, rrc);
    const float ut = x.a1 * vt + x.a2 * tpre + x.b1 * dtt;
    x.t_pre[o] = vt;
    x.ut[o] = pen_div_m<float>(T(i, j, k) + ut, eta, cv1, dt);
    if (x.lf.fx) mark_dirty_lines(x.lf, g, i, j, k, mk);
  }
}


}  // namespace vk

using namespace vk;

extern "C" int vk_time_step_mixed(double* u, double* v, double* w, double* p, float* t,
                                  float* u_pre, float* v_pre, float* w_pre, float* t_pre,
                                  float* u_ta, float* v_ta, float* w_ta, double* u_temp,
                                  double* v_temp, double* w_temp, double* divu, float* cellvalue,
                                  const float* rhocp, const float* thc, const double* dx,
                                  const double* dy, const double* dz, const double* dx12,
                                  const double* dy12, const double* dz12, const int* nx,
                                  const int* ny, const int* nz, const double* dt,
                                  const double* re, const double* sc, const double* eta,
                                  const double* xl, const double* zl, const int* l) {
  VK_REQUIRE_INIT();
  Geom g(*nx, *ny, *nz);
  if (g.nx < 4 || g.ny < 4 || g.nz < 4) return fail(VK_ERR_BAD_SIZE, "time_step_mixed: n >= 4");
  if (!g.fits32()) return fail(VK_ERR_BAD_SIZE, "time_step_mixed: grid too large");
  if (int frc = fft_sizes_ok(g)) return frc;
  const size_t fd = g.cells() * sizeof(double), ff = g.cells() * sizeof(float);
  VK_RESOLVE(du, double, u, fd);
  VK_RESOLVE(dv, double, v, fd);
  VK_RESOLVE(dw, double, w, fd);
  VK_RESOLVE(dp, double, p, fd);
  VK_RESOLVE(dtt, float, t, ff);
  VK_RESOLVE(dup, float, u_pre, ff);
  VK_RESOLVE(dvp, float, v_pre, ff);
  VK_RESOLVE(dwp, float, w_pre, ff);
  VK_RESOLVE(dtp, float, t_pre, ff);
  VK_RESOLVE(duta, float, u_ta, ff);
  VK_RESOLVE(dvta, float, v_ta, ff);
  VK_RESOLVE(dwta, float, w_ta, ff);
  VK_RESOLVE(dut, double, u_temp, fd);
  VK_RESOLVE(dvt, double, v_temp, fd);
  VK_RESOLVE(dwt, double, w_temp, fd);
  VK_RESOLVE(ddiv, double, divu, fd);
  VK_RESOLVE(dcv, float, cellvalue, g.interior() * 4 * sizeof(float));
  VK_RESOLVE(drc, const float, rhocp, ff);
  VK_RESOLVE(dk, const float, thc, ff);
  VK_RESOLVE(ddx, const double, dx, (g.nx + 2) * sizeof(double));
  VK_RESOLVE(ddy, const double, dy, (g.ny + 2) * sizeof(double));
  VK_RESOLVE(ddz, const double, dz, (g.nz + 2) * sizeof(double));
  VK_RESOLVE(ddx12, const double, dx12, (g.nx + 2) * sizeof(double));
  VK_RESOLVE(ddy12, const double, dy12, (g.ny + 2) * sizeof(double));
  VK_RESOLVE(ddz12, const double, dz12, (g.nz + 2) * sizeof(double));
  Metrics m{ddx, ddy, ddz, ddx12, ddy12, ddz12};
  float* twork = static_cast<float*>(workspace(kWorkGrad, ff));
  if (!twork) return VK_ERR_HIP;
  hipStream_t s = rt().stream;

  launch_set_bc<float>(g, du, dv, dw, dp, dtt, s);
  if (*l == 1) launch_cv_faces<float>(g, dcv, s);
  FMet f;
  int rc = make_fmetrics(g, m, f, s);
  if (rc) return rc;
  RMetrics r;
  rc = make_rmetrics(g, m, r, s);
  if (rc) return rc;
  LineFlags lf{};
  rc = line_flags_clear(g, lf, s);
  if (rc) return rc;
  ExplicitMixedArgs x{};
  x.lf = lf;
  x.u = du; x.v = dv; x.w = dw; x.t = dtt; x.rc = drc; x.thc = dk;
  x.u_ta = duta; x.v_ta = dvta; x.w_ta = dwta; x.cv = dcv;
  x.u_pre = dup; x.v_pre = dvp; x.w_pre = dwp; x.t_pre = dtp;
  x.ua = dut; x.va = dvt; x.wa = dwt; x.ut = twork;
  x.dt = (float)*dt; x.eta = (float)*eta;
  x.dtd = *dt; x.etad = *eta;
  x.a1 = (*l == 1) ? 1.0f : 1.5f;
  x.a2 = (*l == 1) ? 0.0f : -0.5f;
  x.b1 = 1.0f;
  x.c0 = (float)(*dt / *re / 2.0);
  x.cd = (float)(*dt / (2.0 * *re * *sc));
  const double beta_u = *dt / *re / 2.0, beta_t = *dt / *re / *sc / 2.0, dteta = *dt * *eta;
  {
    // the explicit phase of u, v, w and T in one kernel (the velocity read
    // once); every later kernel that writes u, v, w runs after it
    ProfScope ps("explicit_mixed", s);
    k_explicit_mixed<<<interior_grid(g), interior_block(), 0, s>>>(g, f, x);
  }
  // T's sweeps and halo on auxiliary stream 2 beside the velocity sweeps
  // and MKP; the y sweep of T writes T's interior
  AuxFork ft(s, 4u);
  if (ft.err) return ft.err;
  rc = launch_adi_temperature<float>(g, m, r, twork, dcv, drc, dk, beta_t, dteta, ft.aux(2),
                                     dtt, &lf);
  if (rc) return rc;
  launch_set_bc<float>(g, nullptr, nullptr, nullptr, nullptr, dtt, ft.aux(2));
  // u and v on auxiliary streams 0 and 1, w here
  {
    AuxFork fv(s, 3u);
    if (fv.err) return fv.err;
    for (int c = 0; c < 2; c++) {
      rc = launch_adi_velocity<float>(g, m, r, c == 0 ? dut : dvt, dcv, c, beta_u, dteta,
                                      fv.aux(c), &lf);
      if (rc) return rc;
    }
    rc = launch_adi_velocity<float>(g, m, r, dwt, dcv, 2, beta_u, dteta, s, &lf);
    if (rc) return rc;
  }
  launch_set_bc<float>(g, dut, dvt, dwt, dp, nullptr, s);
  MkpOutMixed o{dp, duta, dvta, dwta, ddiv, du, dv, dw};
  rc = launch_mkp_mixed(g, m, dut, dvt, dwt, *dt, *xl, *zl, o, s);
  if (rc) return rc;
  rc = ft.join(2);
  if (rc) return rc;
  VK_LAUNCH_CHECK("time_step_mixed");
  return VK_OK;
}
