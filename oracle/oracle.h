/*
 * oracle.h -- CPU restatement of the vtkmakeCompare time-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, the
 * Fortran host, the Python host mirror) may link, load or call this code.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
 * it, and only as the checker / the CPU baseline.
 *
 * Every function follows one reference routine (file:line given beside it),
 * keeps its argument order, and performs the same floating-point operations
 * in the same order (build with -ffp-contract=off).  Arrays follow the
 * reference layout: halo fields are (0:nx+1,0:ny+1,0:nz+1) column-major
 * (x fastest), cellvalue is (1:nx,1:ny,1:nz,1:4), spectra are complex
 * (0:nx/2,0:ny+1,0:nz+1) stored as interleaved (re,im) doubles.
 * Scalars are passed by value.
 *
 * Defined semantics where the reference is undefined (the library defines
 * them the same way):
 *  - harmonic-mean conductivity ghosts of the T line solves
 *    (nse_dominik.f:1755,1760,1814,1819 read llx(0), llx(n+1), lly(0),
 *    lly(ny+1), which parallel_Tx/Ty/Tz never fill): x and z wrap
 *    periodically, y copies the wall-adjacent value;
 *  - set_adj_NS_scalar reads an uninitialised local T_final
 *    (adjoint.f:190); here T_final is an extra trailing argument;
 *  - MKP's automatic PHI array starts zeroed, so the corner halos PHIBC
 *    leaves unset are 0 instead of stack garbage.
 */
#ifndef VK_ORACLE_H
#define VK_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* No reference routine: phi0 from a union of ellipsoids, the analytic
 * stand-in for the structure.dat that READGRID3D reads
 * (set_levelset.f:202-214).  shapes holds nshapes records (cx, cy, cz, a,
 * b, c, qw, qx, qy, qz); phi0 = max over shapes of
 * min(a,b,c) (1 - |R^T d / (a,b,c)|), d the minimum-image offset in x and
 * z, R the rotation of the normalised quaternion; -1 with no shape; halo 0.
 * Pinned by closed forms (tests/test_oracle_math.py): a sphere gives
 * r - |d|, surface points of a rotated ellipsoid give 0, and the sign
 * tells inside from outside. */
void o_levelset_shapes(int nx, int ny, int nz, const double *x, const double *y,
                       const double *z, double xl, double zl, int nshapes,
                       const double *shapes, double *phi0);

/* The iterative Poisson solver the reference keeps but does not call.
 * nse_dominik.f:2136-2267 (linebyline) with implicit_px/py/pz
 * (:2270-2457): one line Gauss-Seidel iteration on p for
 * L p = div(u_temp)/dt -- x lines in (k, j) order, then y lines in (k, i)
 * order, then z lines in (j, i) order, each line solved with its last
 * row's metrics for every row -- then zero-gradient halos on all six faces
 * and p shifted so that p(1,1,1) = 0.  The reference declares fpy and fpz
 * with nx+2 elements but fills ny+2 and nz+2 of them (undefined past nx+1);
 * here they are as long as they need to be. */
void o_linebyline(int nx, int ny, int nz, const double *u_temp,
                  const double *v_temp, const double *w_temp, const double *dx,
                  const double *dy, const double *dz, const double *dx12,
                  const double *dy12, const double *dz12, double dt, double *p);
/* nse_dominik.f:1955-2021 (solve_p): linebyline until the largest residual
 * |L p dt - div u_temp| is below 1e-5 or max_iter iterations (the
 * reference's nnn = 100000); res is the last interior point's residual,
 * res_int and res_var the volume integrals of the residual and of its
 * square; *iters the iterations run. */
void o_solve_p(double *p, const double *u_temp, const double *v_temp,
               const double *w_temp, double dt, const double *dx,
               const double *dy, const double *dz, const double *dx12,
               const double *dy12, const double *dz12, int nx, int ny, int nz,
               double *res, double *res_max, double *res_int, double *res_var,
               int max_iter, int *iters);

/* nse_dominik.f:553-607 */
void o_set_c(int nx, int ny, int nz, double *x, double *x12, double *y,
             double *y12, double *z, double *z12, double *dx, double *dx12,
             double *dy, double *dy12, double *dz, double *dz12, double xl,
             double yl, double zl);
/* nse_dominik.f:611-753 */
void o_set_bc(double *u, double *v, double *w, double *p, int nx, int ny,
              int nz, double *t);
/* nse_dominik.f:757-831 */
void o_set_convect(const double *u, double *u_v, const double *v, double *v_v,
                   const double *w, double *w_v, double dt, int nx, int ny,
                   int nz, const double *dx, const double *dy, const double *dz,
                   const double *dx12, const double *dy12, const double *dz12);
/* nse_dominik.f:834-884 */
void o_set_convection_for_scalar(const double *u, const double *v,
                                 const double *w, double dt, int nx, int ny,
                                 int nz, const double *dx, const double *dy,
                                 const double *dz, const double *dx12,
                                 const double *dy12, const double *dz12,
                                 const double *t, double *u_vt,
                                 const double *rhocp);
/* nse_dominik.f:888-945 */
void o_set_conduction(const double *u, double *u_d, const double *v,
                      double *v_d, const double *w, double *w_d, double re,
                      double dt, int nx, int ny, int nz, const double *dx,
                      const double *dy, const double *dz, const double *dx12,
                      const double *dy12, const double *dz12);
/* nse_dominik.f:950-1004 */
void o_set_diffusion_for_scalar(double re, double dt, int nx, int ny, int nz,
                                const double *dx, const double *dy,
                                const double *dz, const double *dx12,
                                const double *dy12, const double *dz12,
                                const double *t, double *u_dt, double sc,
                                const double *thc, const double *rhocp);
/* nse_dominik.f:1011-1129 */
void o_set_immersed_boundary(double eta, int nx, int ny, int nz, double dt,
                             double *u_l, double *v_l, double *w_l,
                             const double *dx, const double *dy,
                             const double *dz, double *t_l, double *cellvalue,
                             const double *u_ta, const double *v_ta,
                             const double *w_ta, int l, double *cons_cell);
/* nse_dominik.f:1132-1286 */
void o_set_explicit(const double *u_v, double *u_pre, const double *u_d,
                    double *u_a, const double *v_v, double *v_pre,
                    const double *v_d, double *v_a, const double *w_v,
                    double *w_pre, const double *w_d, double *w_a,
                    const double *u_l, const double *v_l, const double *w_l,
                    int nx, int ny, int nz, int l);
/* nse_dominik.f:1291-1341 */
void o_set_explicit_for_tempreture(const double *u_vt, const double *u_dt,
                                   double *u_t, double *t_pre,
                                   const double *t_l, int nx, int ny, int nz,
                                   int l);
/* nse_dominik.f:387-396 (main-loop temperature penalisation divide) */
void o_t_divide(double *u_t, const double *t, const double *cellvalue,
                double eta, double dt, int nx, int ny, int nz);
/* nse_dominik.f:1344-1461 */
void o_solve_u_temp(double *u_a, double *v_a, double *w_a, int nx, int ny,
                    int nz, const double *dx, const double *dy,
                    const double *dz, const double *dx12, const double *dy12,
                    const double *dz12, double re, double dt, double eta,
                    const double *cellvalue, const double *u, const double *v,
                    const double *w);
/* nse_dominik.f:1465-1501 */
void o_solve_t(double *u_t, int nx, int ny, int nz, const double *dx,
               const double *dy, const double *dz, const double *dx12,
               const double *dy12, const double *dz12, double re, double sc,
               double dt, const double *rhocp, const double *thc,
               const double *cellvalue, double eta);

/* line solvers (one line each) */
/* nse_dominik.f:1505-1577 */
void o_implicit_uy(double *suy, int ny, const double *dy, const double *dy12,
                   double re, double dt, double ft, double fb, double eta,
                   const double *cuy);
/* nse_dominik.f:1583-1648 */
void o_implicit_vy(double *svy, int ny, const double *dy, const double *dy12,
                   double re, double dt, double ft, double fb, double eta,
                   const double *cvy);
/* nse_dominik.f:1725-1781 (cux indexed 0..nx+1 as in the reference) */
void o_implicit_tx(double *stx, int nx, const double *dx, const double *dx12,
                   double re, double sc, double dt, const double *llx,
                   const double *rrx, double ft, double fb, double eta,
                   const double *cux);
/* nse_dominik.f:1784-1870 */
void o_implicit_ty(double *sty, int ny, const double *dy, const double *dy12,
                   double re, double sc, double dt, const double *lly,
                   const double *rry, double ft, double fb, const double *cuy,
                   double eta);
/* periodic_implicit.f:1-44 (cux indexed 1..nx: pass pointer to element 1) */
void o_implicit_solver(double *sux, int nx, const double *dx,
                       const double *dx12, double re, double dt, double ft,
                       double fb, double eta, const double *cux);
/* periodic_implicit.f:47-64 */
void o_tdma(const double *a, const double *b, const double *c, const double *r,
            double *u, int n);
/* periodic_implicit.f:67-93 */
void o_cyclic(const double *a, const double *b, const double *c, double alpha,
              double beta, const double *r, double *x, int n);
/* tridiag.f:43-58; a is (nn,3) column-major, b complex (re,im pairs) */
void o_ctridiag(double *a, double *b, int nn);

/* pressure projection */
/* mkp.f:9-220 */
void o_mkp(double *p, const double *u_temp, const double *v_temp,
           const double *w_temp, double dt, const double *dx,
           const double *dy, const double *dz, const double *dx12,
           const double *dy12, const double *dz12, double re, int nx, int ny,
           int nz, double *u_ta, double *v_ta, double *w_ta, double *u_ep2,
           double *v_ep2, double *w_ep2, double *divu, double *u, double *v,
           double *w, double xl, double zl);
/* p2s_fftw2d.f:5-114 (index/scaling conventions; the FFT is this file's) */
void o_p2s(const double *fp, double *fpt, int nx, int ny, int nz);
/* p2s_fftw2d.f:116-223 */
void o_s2p(const double *fpt, double *fp, int nx, int ny, int nz);
/* mkp.f:223-260 (ac is (1:ny,3), kxz is (0:nx/2,0:nz)) */
void o_spn(int nx, int ny, int nz, const double *ac, const double *kxz,
           double *phit);
/* mkp.f:263-290 */
void o_sp0b(int nx, int ny, int nz, const double *ac, double *phit);
/* mkp.f:293-320 */
void o_phibc(int nx, int ny, int nz, double *phi);
/* mkp.f:54-84: wavenumber table KXZ(0:nx/2,0:nz) and AC(1:ny,3) */
void o_mkp_coefficients(int nx, int ny, int nz, const double *dx,
                        const double *dy, const double *dz,
                        const double *dy12, double xl, double zl, double *kxz,
                        double *ac);
/* complex DFT used by o_p2s/o_s2p: sign=-1 forward, +1 backward,
   unnormalised; data interleaved (re,im) with the given element stride */
void o_fft(double *data, int n, int stride, int sign);

/* level set and adjoint operators */
/* set_levelset.f:1-199 (VTK dumps and the printed residual omitted) */
void o_set_levelset(int nx, int ny, int nz, const double *dx,
                    const double *dx12, const double *dy12, const double *dz12,
                    const double *phi0, double *phi, const double *x,
                    const double *y, const double *z, double *cellvalue,
                    int iterations);
/* adjoint.f:1-126 */
void o_set_adj_convect(const double *u, double *u_v, const double *v,
                       double *v_v, const double *w, double *w_v, double re,
                       double dt, int nx, int ny, int nz, const double *dx,
                       const double *dy, const double *dz, const double *dx12,
                       const double *dy12, const double *dz12,
                       const double *u_final, const double *v_final,
                       const double *w_final);
/* adjoint.f:128-181 */
void o_set_adj_convection_for_scalar(
    const double *u, const double *v, const double *w, double re, double dt,
    int nx, int ny, int nz, const double *dx, const double *dy,
    const double *dz, const double *dx12, const double *dy12,
    const double *dz12, const double *t, double *u_vt, const double *rhocp,
    const double *u_final, const double *v_final, const double *w_final);
/* adjoint.f:185-211 (+ T_final argument) */
void o_set_adj_ns_scalar(double dt, int nx, int ny, int nz,
                         const double *dx12, const double *dy12,
                         const double *dz12, const double *t, double *u_s,
                         double *v_s, double *w_s, const double *u_final,
                         const double *v_final, const double *w_final,
                         const double *t_final);

/* One pass of the main time loop body, nse_dominik.f:290-448, in the
   reference's call order.  The report/timer/output branches and the
   T_before diagnostic copies are not part of the step. */
typedef struct {
  int nx, ny, nz;
  double dt, re, sc, eta, xl, zl;
  double *dx, *dy, *dz, *dx12, *dy12, *dz12;
  double *u, *v, *w, *p, *t;
  double *u_v, *v_v, *w_v, *u_d, *v_d, *w_d, *u_a, *v_a, *w_a;
  double *u_temp, *v_temp, *w_temp, *u_pre, *v_pre, *w_pre;
  double *u_ta, *v_ta, *w_ta, *u_l, *v_l, *w_l;
  double *u_vt, *u_dt, *u_t, *t_pre, *t_l;
  double *rhocp, *thc, *cellvalue;
  double *u_ep2, *v_ep2, *w_ep2, *divu;
} o_state;
void o_time_step(o_state *s, int l);

/* ------------------------------------------------------- diagnostics */
/* CHECK_COURANT nse_dominik.f:2519-2601: the largest of u dt/dx(i),
   v dt/dy(j), w dt/dz(k) over the interior (signed, from 0), which the
   reference prints; here returned in *courant. */
void o_check_courant(double dt, const double *dx, const double *dy, const double *dz, int nx,
                     int ny, int nz, const double *u, const double *v, const double *w,
                     double *courant);
/* MONITOR nse_dominik.f:2797-2872, same argument order (the scalars by
   pointer as in the reference); loc[3] receives the (i, j, k) of the
   largest |T - T_before| that the reference prints as xx, yy, zz. */
void o_monitor(int nx, int ny, int nz, int l, const double *t, const double *dx,
               const double *dy, const double *dz, double *dif, double *t_all, double *t_dif,
               const double *t_before, const double *u, const double *v, const double *w,
               double *vel_dif, double *u_dif, double *v_dif, double *w_dif,
               double *u_int_before, double *v_int_before, double *w_int_before, double *t_b,
               double dt, int *loc);
/* CHECK_CONVERGENCE nse_dominik.f:2875-2892: the error sum that the
   reference appends to timely_convergence.dat. */
void o_check_convergence(const double *y, double yl, const double *u, double re, int nx,
                         int ny, int nz, double *error_sum);

#ifdef __cplusplus
}
#endif
#endif
