// vk_step.hip -- one pass of the main-loop body (nse_dominik.f:290-448)
// with its pointwise phases fused:
//
//   SET_BC(u,v,w,p,T)                                     halo kernel
//   A: SET_CONVECT, SET_CONVECTION_FOR_SCALAR, SET_CONDUCTION,
//      SET_DIFFUSION_FOR_SCALAR, SET_IMMERSED_BOUNDARY, SET_EXPLICIT,
//      set_explicit_for_tempreture, the T divide (:387-396) and the
//      penalisation divides of solve_u_temp              one kernel
//   B-D: ADI sweeps of solve_u_temp (y,z,x) and solve_t (y,x,z)
//   SET_BC(u_temp,v_temp,w_temp,p,T)                      halo kernel
//   MKP                                                   FFT/solve/project
//
// Intermediate arrays of the reference (u_v, u_d, u_l, u_a, u_vT, u_dT,
// T_l, u_T) never reach HBM: phase A writes the divided right-hand sides
// straight into u_temp/v_temp/w_temp and a temperature work field, and
// the sweeps solve in place.  Every value written is computed with the
// reference's operation order.
#include "vk_kernels.h"
#include "vk_runtime.h"
#include "vk_stencil_ops.h"

namespace vk {

// x / (1 + eta cv dt); where cv is 0 the divisor is exactly 1 and the
// quotient exactly x, so no branch is needed to keep fluid cells exact
__device__ inline double pen_div(double x, double eta, double cv, double dt) {
  return x / (1.0 + eta * cv * dt);
}

// u, v, w, then T, in one pass: T takes the velocity at the point and its
// lower neighbours from the stencils loaded for u, v, w, so the velocity is
// read once, and T's loads are issued before the velocity's stores, so
// they are in flight while those go out.
__global__ __launch_bounds__(256) void k_explicit_fused(Geom g, Metrics m, RMetrics r,
                                                        ExplicitArgs x) {
  int bx, by, bz;
  xcd_block(bx, by, bz);
  int i = 1 + bx * blockDim.x + threadIdx.x;
  int j = 1 + by * blockDim.y + threadIdx.y;
  int k = 1 + bz;
  if (i > g.nx || j > g.ny) return;
  const long long o = g.at(i, j, k);
  const double dt = x.dt, eta = x.eta;
  {
    // the point's cellvalue slots, read before any store (the outputs may
    // alias them as far as the compiler knows)
    const double cv2 = x.cv[g.cv(i, j, k, 2)], cv3 = x.cv[g.cv(i, j, k, 3)],
                 cv4 = x.cv[g.cv(i, j, k, 4)];
    const double cv1 = x.cv[g.cv(i, j, k, 1)];
    // the three stencils and the history, every load issued before any
    // arithmetic (left to itself the compiler issues a few loads, waits,
    // computes and issues the next few); the values and the operations on
    // them are those of the F3 views
    Nb27<double, double> U, V, W;
    load_velocity_nbhd(U, V, W, x.u, x.v, x.w, g, i, j, k);
    const double upre = x.u_pre[o], vpre = x.v_pre[o], wpre = x.w_pre[o];
    PtMet<double> M;
    M.load(m, r, i, j, k);
    __builtin_amdgcn_sched_barrier(0);
    unsigned mk = x.lf.fx ? dirty_bits(0.0, cv2, cv3, cv4, 1.0, 1.0) : 0u;
    // u_ta enters only through eta cv u_ta dt: where the face value is 0
    // the term is 0 and u_ta is not read (in fluid, most of the grid, the
    // three fields' traffic is saved)
    double uta = 0.0, vta = 0.0, wta = 0.0;
    if (cv2 != 0.0) uta = x.u_ta[o];
    if (cv3 != 0.0) vta = x.v_ta[o];
    if (cv4 != 0.0) wta = x.w_ta[o];
    // u, v, w: SET_CONVECT, SET_CONDUCTION, u_l, SET_EXPLICIT and
    // solve_u_temp's divide.  Every value is computed before the first
    // store, so the loads of all three stencils can be in flight together
    // (a store would otherwise fence the later loads: the outputs may
    // alias the inputs as far as the compiler knows).
    const double uv = conv_u_r(U, V, W, M, M, i, j, k, dt);
    const double ud = cond_u_r(U, M, i, j, k, x.c0);
    const double ul = eta * cv2 * uta * dt;
    const double ua = x.a1 * uv + x.a2 * upre + x.b1 * ud + x.b1 * ul;
    const double vv = conv_v_r(U, V, W, M, M, i, j, k, dt);
    const double vd = cond_v_r(V, M, i, j, k, x.c0);
    const double vl = eta * cv3 * vta * dt;
    const double va = x.a1 * vv + x.a2 * vpre + x.b1 * vd + x.b1 * vl;
    const double wv = conv_w_r(U, V, W, M, M, i, j, k, dt);
    const double wd = cond_w_r(W, M, i, j, k, x.c0);
    const double wl = eta * cv4 * wta * dt;
    const double wa = x.a1 * wv + x.a2 * wpre + x.b1 * wd + x.b1 * wl;
    const double un = pen_div(U(i, j, k) + ua, eta, cv2, dt);
    const double vn = pen_div(V(i, j, k) + va, eta, cv3, dt);
    const double wn = pen_div(W(i, j, k) + wa, eta, cv4, dt);
    // T's 7-point stencils of T, rhoCp and thc
    Nb27<double, double> T, RC, K;
    load_scalar_fields(T, RC, K, x.t, x.rc, x.thc, g, i, j, k);
    const double tpre = x.t_pre[o];
    x.u_pre[o] = uv;
    x.ua[o] = un;
    x.v_pre[o] = vv;
    x.va[o] = vn;
    // SET_EXPLICIT's reflection v_a(i,ny+1,k) = -v_a(i,ny,k) survives into
    // v_temp (no later routine writes that plane of v)
    if (j == g.ny) x.va[g.at(i, g.ny + 1, k)] = -1.0 * va;
    x.w_pre[o] = wv;
    x.wa[o] = wn;
    // T: convection, diffusion, T_l = 0, combine, history, main-loop
    // divide, with the velocity values of U, V, W
    __builtin_amdgcn_sched_barrier(0);
    const double rc0 = RC(i, j, k);
    if (x.lf.fx) mk |= dirty_bits(cv1, 0.0, 0.0, 0.0, rc0, K(i, j, k));
    double rrc = 1.0 / rc0;
    double vt = conv_scalar_r(U, V, W, T, RC, M, M, i, j, k, dt, rrc);
    double dtt = diff_scalar_r(T, K, M, M, i, j, k, x.cd, rrc);
    double ut = x.a1 * vt + x.a2 * tpre + x.b1 * dtt;
    x.t_pre[o] = vt;
    x.ut[o] = pen_div(T(i, j, k) + ut, eta, cv1, dt);
    if (x.lf.fx) mark_dirty_lines(x.lf, g, i, j, k, mk);
  }
}


__global__ void k_fill_u32(unsigned* p, long long n) {
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (long long)gridDim.x * blockDim.x)
    p[q] = 0u;
}

int line_flags_clear(const Geom& g, LineFlags& lf, hipStream_t s) {
  const long long nxl = (long long)g.ny * g.nz, nyl = (long long)g.nx * g.nz,
                  nzl = (long long)g.nx * g.ny, n = nxl + nyl + nzl;
  unsigned* w = static_cast<unsigned*>(workspace(kWorkFlags, (size_t)n * sizeof(unsigned)));
  if (!w) return VK_ERR_HIP;
  lf.fx = w;
  lf.fy = w + nxl;
  lf.fz = w + nxl + nyl;
  ProfScope ps("line_flags", s);
  const long long nb = (n + 255) / 256;
  k_fill_u32<<<(unsigned)(nb < 1024 ? nb : 1024), 256, 0, s>>>(w, n);
  return VK_OK;
}

void launch_explicit_fused(const Geom& g, const Metrics& m, const RMetrics& r,
                           const ExplicitArgs& x, hipStream_t s) {
  ProfScope ps("explicit_fused", s);
  k_explicit_fused<<<interior_grid(g), interior_block(), 0, s>>>(g, m, r, x);
}

ExplicitArgs explicit_args(double dt, double re, double sc, double eta, int l) {
  ExplicitArgs x{};
  x.dt = dt; x.re = re; x.sc = sc; x.eta = eta;
  if (l == 1) {
    x.a1 = 1.0; x.a2 = 0.0; x.b1 = 1.0;
  } else {
    x.a1 = 1.5; x.a2 = -0.5; x.b1 = 1.0;
  }
  x.c0 = dt / re / 2.0;
  x.cd = dt / (2.0 * re * sc);
  return x;
}

}  // namespace vk

using namespace vk;

extern "C" int vk_time_step(double* u, double* v, double* w, double* p, double* t,
                            double* u_pre, double* v_pre, double* w_pre, double* t_pre,
                            double* u_ta, double* v_ta, double* w_ta, double* u_temp,
                            double* v_temp, double* w_temp, double* divu, double* cellvalue,
                            const double* rhocp, const double* thc, double* u_ep2,
                            double* v_ep2, double* w_ep2, const double* dx, const double* dy,
                            const double* dz, const double* dx12, const double* dy12,
                            const double* dz12, const int* nx, const int* ny, const int* nz,
                            const double* dt, const double* re, const double* sc,
                            const double* eta, const double* xl, const double* zl,
                            const int* l) {
  VK_REQUIRE_INIT();
  Geom g(*nx, *ny, *nz);
  if (g.nx < 4 || g.ny < 4 || g.nz < 4) return fail(VK_ERR_BAD_SIZE, "time_step: n >= 4");
  if (!g.fits32()) return fail(VK_ERR_BAD_SIZE, "time_step: grid too large");
  if (int frc = fft_sizes_ok(g)) return frc;
  const size_t fb = g.cells() * sizeof(double);
  VK_RESOLVE(du, double, u, fb);
  VK_RESOLVE(dv, double, v, fb);
  VK_RESOLVE(dw, double, w, fb);
  VK_RESOLVE(dp, double, p, fb);
  VK_RESOLVE(dtt, double, t, fb);
  VK_RESOLVE(dup, double, u_pre, fb);
  VK_RESOLVE(dvp, double, v_pre, fb);
  VK_RESOLVE(dwp, double, w_pre, fb);
  VK_RESOLVE(dtp, double, t_pre, fb);
  VK_RESOLVE(duta, double, u_ta, fb);
  VK_RESOLVE(dvta, double, v_ta, fb);
  VK_RESOLVE(dwta, double, w_ta, fb);
  VK_RESOLVE(dut, double, u_temp, fb);
  VK_RESOLVE(dvt, double, v_temp, fb);
  VK_RESOLVE(dwt, double, w_temp, fb);
  VK_RESOLVE(ddiv, double, divu, fb);
  VK_RESOLVE(dcv, double, cellvalue, g.interior() * 4 * sizeof(double));
  VK_RESOLVE(drc, const double, rhocp, fb);
  VK_RESOLVE(dk, const double, thc, fb);
  double *de1 = nullptr, *de2 = nullptr, *de3 = nullptr;
  if (u_ep2 && v_ep2 && w_ep2) {
    de1 = static_cast<double*>(resolve(u_ep2, fb, "u_ep2"));
    de2 = static_cast<double*>(resolve(v_ep2, fb, "v_ep2"));
    de3 = static_cast<double*>(resolve(w_ep2, fb, "w_ep2"));
    if (!de1 || !de2 || !de3) return last_status();
  }
  VK_RESOLVE(ddx, const double, dx, (g.nx + 2) * sizeof(double));
  VK_RESOLVE(ddy, const double, dy, (g.ny + 2) * sizeof(double));
  VK_RESOLVE(ddz, const double, dz, (g.nz + 2) * sizeof(double));
  VK_RESOLVE(ddx12, const double, dx12, (g.nx + 2) * sizeof(double));
  VK_RESOLVE(ddy12, const double, dy12, (g.ny + 2) * sizeof(double));
  VK_RESOLVE(ddz12, const double, dz12, (g.nz + 2) * sizeof(double));
  Metrics m{ddx, ddy, ddz, ddx12, ddy12, ddz12};
  double* twork = static_cast<double*>(workspace(kWorkGrad, fb));
  if (!twork) return VK_ERR_HIP;
  hipStream_t s = rt().stream;

  launch_set_bc(g, du, dv, dw, dp, dtt, s);
  if (*l == 1) launch_cv_faces(g, dcv, s);
  ExplicitArgs x = explicit_args(*dt, *re, *sc, *eta, *l);
  x.u = du; x.v = dv; x.w = dw; x.t = dtt; x.rc = drc; x.thc = dk;
  x.u_ta = duta; x.v_ta = dvta; x.w_ta = dwta; x.cv = dcv;
  x.u_pre = dup; x.v_pre = dvp; x.w_pre = dwp; x.t_pre = dtp;
  x.ua = dut; x.va = dvt; x.wa = dwt; x.ut = twork;
  RMetrics r;
  int rrc = make_rmetrics(g, m, r, s);
  if (rrc) return rrc;
  // the explicit phase also marks the lines with solid, rhoCp or thc
  // points, which the sweeps then send to their general path
  LineFlags lf{};
  rrc = line_flags_clear(g, lf, s);
  if (rrc) return rrc;
  x.lf = lf;
  const double beta_u = *dt / *re / 2.0, beta_t = *dt / *re / *sc / 2.0, dteta = *dt * *eta;
  // the explicit phase of u, v, w and T in one kernel (the velocity read
  // once); every later kernel that writes u, v, w runs after it
  launch_explicit_fused(g, m, r, x, s);
  // The temperature's sweeps and halo run on auxiliary stream 2 beside the
  // velocity sweeps and MKP.  T = T work after the sweeps
  // (nse_dominik.f:429-437): the y sweep writes T's interior and the x and
  // z sweeps run there.
  AuxFork ft(s, 4u);
  if (ft.err) return ft.err;
  rrc = launch_adi_temperature(g, m, r, twork, dcv, drc, dk, beta_t, dteta, ft.aux(2), dtt, &lf);
  if (rrc) return rrc;
  launch_set_bc<double>(g, nullptr, nullptr, nullptr, nullptr, dtt, ft.aux(2));
  // the velocity components' sweeps are independent: u and v on auxiliary
  // streams 0 and 1, w here
  {
    AuxFork fv(s, 3u);
    if (fv.err) return fv.err;
    for (int c = 0; c < 2; c++) {
      rrc = launch_adi_velocity(g, m, r, c == 0 ? dut : dvt, dcv, c, beta_u, dteta, fv.aux(c),
                                &lf);
      if (rrc) return rrc;
    }
    rrc = launch_adi_velocity(g, m, r, dwt, dcv, 2, beta_u, dteta, s, &lf);
    if (rrc) return rrc;
  }
  launch_set_bc<double>(g, dut, dvt, dwt, dp, nullptr, s);
  MkpOut o{dp, duta, dvta, dwta, de1, de2, de3, ddiv, du, dv, dw};
  int rc = launch_mkp(g, m, dut, dvt, dwt, *dt, *xl, *zl, o, s);
  if (rc) return rc;
  rc = ft.join(2);
  if (rc) return rc;
  VK_LAUNCH_CHECK("time_step");
  return VK_OK;
}
