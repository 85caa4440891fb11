/*
 * vtkgpu.h -- C ABI of the MI355X (gfx950) library for the vtkmakeCompare
 * time-step hot path.
 *
 * Each reference routine on the path has one entry point with the
 * reference's argument list and order.  Every argument is passed by
 * pointer, as a Fortran caller passes it (`bind(C)` interfaces in
 * vtkmakecompare_amd/host/vk_capi.f90), so a Fortran host keeps its call
 * sites.  Array arguments may be
 *   - host arrays registered with vk_register (the call then works on the
 *     device mirror; vk_upload / vk_download move data), or
 *   - device pointers (vk_malloc or any HIP allocation on this device).
 * Arrays keep the reference layout: halo fields (0:nx+1,0:ny+1,0:nz+1)
 * column-major doubles, cellvalue (1:nx,1:ny,1:nz,1:4), spectra
 * (0:nx/2,0:ny+1,0:nz+1) complex double (re,im interleaved).
 *
 * All entry points return a status (VK_OK = 0); vk_last_error() gives the
 * message.  Work is queued on the library's stream; vk_download and
 * vk_synchronize wait for it.  Calls come from one host thread.
 */
#ifndef VTKGPU_H
#define VTKGPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  VK_OK = 0,
  VK_ERR_NOT_INITIALISED = 1,
  VK_ERR_NOT_RESIDENT = 2, /* array neither registered nor device memory */
  VK_ERR_BAD_SIZE = 3,     /* grid size the kernels do not support */
  VK_ERR_HIP = 4,          /* HIP runtime error */
  VK_ERR_ARGUMENT = 5,     /* invalid argument */
  VK_ERR_NONFINITE = 6,    /* a field check found NaN or Inf */
  VK_ERR_GUARD = 7         /* vk_guard_check found a guard zone written */
};

/* ------------------------------------------------------------ runtime */
/* No reference counterpart: the reference runs in host memory. */
int vk_init(const int *device);
int vk_finalize(void);
const char *vk_last_error(void);
int vk_register(const void *host, const size_t *bytes);
int vk_unregister(const void *host);
int vk_upload(const void *host);
int vk_download(void *host);
int vk_synchronize(void);
int vk_malloc(void **dptr, const size_t *bytes);
int vk_free(void *dptr);
int vk_memcpy_htod(void *dst, const void *src, const size_t *bytes);
int vk_memcpy_dtoh(void *dst, const void *src, const size_t *bytes);
int vk_memset(void *dst, const int *value, const size_t *bytes);
/* Bytes from p to the end of the registered host array or device
   allocation that holds it.  Every entry point makes the same check on
   its arguments and refuses (VK_ERR_ARGUMENT) an array too small for the
   grid it is given. */
int vk_extent(const void *p, size_t *bytes);
/* Out-of-bounds check of the library's own device buffers.  No reference
   counterpart: a check for test suites and debugging.  With VK_GUARD set
   (and not "0") in the environment when vk_init runs, every device
   allocation the library makes (vk_malloc, registered arrays' mirrors,
   workspace, tables, checkpoints) lies between two 64 KiB zones of a byte
   pattern.  vk_guard_check waits for the device, compares every zone with
   the pattern, sets *damaged to the number of zones that changed (a kernel
   wrote past an end of its buffer), refills them, and returns VK_ERR_GUARD
   naming the allocations and offsets when there are any.  Without guard
   zones it sets *damaged to -1 and returns VK_OK. */
int vk_guard_check(int *damaged);

/* Kernel timing on the library stream with HIP events, for benchmarks:
   between vk_profile_begin and vk_profile_end every kernel launch is
   bracketed by an event pair under its kernel's name; vk_profile_entry
   returns the summed duration and the launch count per name.  While
   timing, the fused steps issue the work they otherwise spread over
   several streams on the library stream, one kernel after another, so
   each duration is the kernel's own.  No reference counterpart (the
   reference prints system_clock deltas). */
int vk_profile_begin(void);
int vk_profile_end(void);
int vk_profile_count(int *n);
int vk_profile_entry(const int *index, char *name, const int *name_len,
                     double *total_ms, int *launches);

/* ---------------------------------------------------- boundary halos */
/* SET_BC(U,V,W,P,NX,NY,NZ,T), nse_dominik.f:611-753 */
int vk_set_bc(double *u, double *v, double *w, double *p, const int *nx,
              const int *ny, const int *nz, double *t);

/* ---------------------------------------------------------- stencils */
/* SET_CONVECT, nse_dominik.f:757-831 */
int vk_set_convect(const double *u, double *u_v, const double *v,
                   double *v_v, const double *w, double *w_v,
                   const double *dt, const int *nx, const int *ny,
                   const int *nz, const double *dx, const double *dy,
                   const double *dz, const double *dx12, const double *dy12,
                   const double *dz12);
/* SET_CONVECTION_FOR_SCALAR, nse_dominik.f:834-884 */
int vk_set_convection_for_scalar(const double *u, const double *v,
                                 const double *w, const double *dt,
                                 const int *nx, const int *ny, const int *nz,
                                 const double *dx, const double *dy,
                                 const double *dz, const double *dx12,
                                 const double *dy12, const double *dz12,
                                 const double *t, double *u_vt,
                                 const double *rhocp);
/* SET_CONDUCTION, nse_dominik.f:888-945 */
int vk_set_conduction(const double *u, double *u_d, const double *v,
                      double *v_d, const double *w, double *w_d,
                      const double *re, const double *dt, const int *nx,
                      const int *ny, const int *nz, const double *dx,
                      const double *dy, const double *dz, const double *dx12,
                      const double *dy12, const double *dz12);
/* SET_DIFFUSION_FOR_SCALAR, nse_dominik.f:950-1004 */
int vk_set_diffusion_for_scalar(const double *re, const double *dt,
                                const int *nx, const int *ny, const int *nz,
                                const double *dx, const double *dy,
                                const double *dz, const double *dx12,
                                const double *dy12, const double *dz12,
                                const double *t, double *u_dt,
                                const double *sc, const double *thc,
                                const double *rhocp);
/* SET_IMMERSED_BOUNDARY, nse_dominik.f:1011-1129 */
int vk_set_immersed_boundary(const double *eta, const int *nx, const int *ny,
                             const int *nz, const double *dt, double *u_l,
                             double *v_l, double *w_l, const double *dx,
                             const double *dy, const double *dz, double *t_l,
                             double *cellvalue, const double *u_ta,
                             const double *v_ta, const double *w_ta,
                             const int *l, double *cons_cell);
/* SET_EXPLICIT, nse_dominik.f:1132-1286 */
int vk_set_explicit(const double *u_v, double *u_pre, const double *u_d,
                    double *u_a, const double *v_v, double *v_pre,
                    const double *v_d, double *v_a, const double *w_v,
                    double *w_pre, const double *w_d, double *w_a,
                    const double *u_l, const double *v_l, const double *w_l,
                    const int *nx, const int *ny, const int *nz,
                    const int *l);
/* set_explicit_for_tempreture, nse_dominik.f:1291-1341 */
int vk_set_explicit_for_tempreture(const double *u_vt, const double *u_dt,
                                   double *u_t, double *t_pre,
                                   const double *t_l, const int *nx,
                                   const int *ny, const int *nz,
                                   const int *l);
/* main-loop temperature divide u_T = (T+u_T)/(1+eta cv1 dt),
   nse_dominik.f:387-396 (inline in program main, no routine name) */
int vk_t_divide(double *u_t, const double *t, const double *cellvalue,
                const double *eta, const double *dt, const int *nx,
                const int *ny, const int *nz);

/* ----------------------------------------------------- implicit sweeps */
/* solve_u_temp, nse_dominik.f:1344-1461 (with implicit_uy/vy,
   implicit_solver/cyclic/TDMA and the parallel_* gathers) */
int vk_solve_u_temp(double *u_a, double *v_a, double *w_a, const int *nx,
                    const int *ny, const int *nz, const double *dx,
                    const double *dy, const double *dz, const double *dx12,
                    const double *dy12, const double *dz12, const double *re,
                    const double *dt, const double *eta,
                    const double *cellvalue, const double *u, const double *v,
                    const double *w);
/* solve_t, nse_dominik.f:1465-1501 (implicit_Tx/Ty, parallel_Tx/Ty/Tz).
   The conductivity ghosts the reference leaves undefined
   (nse_dominik.f:1755,1760,1814,1819) are x/z periodic and y wall copies. */
int vk_solve_t(double *u_t, const int *nx, const int *ny, const int *nz,
               const double *dx, const double *dy, const double *dz,
               const double *dx12, const double *dy12, const double *dz12,
               const double *re, const double *sc, const double *dt,
               const double *rhocp, const double *thc,
               const double *cellvalue, const double *eta);

/* ------------------------------------------------- batched line solves
   The reference's line-level solvers, nsys lines per call.  The reference
   solves one line per call from inside its OpenMP sweeps; here each call
   solves a batch, one lane per line, with the reference's elimination in
   its operation order (bit for bit per line).  Every line array is a
   Fortran array X(nsys, 0:n+1) -- the systems fastest -- unless stated;
   per-line scalars of the reference (ft, fb, alpha, beta) are arrays of
   nsys; the metrics (dx, dx12, ...) are shared, n+2 values.  With
   nsys = 1 every array has the reference's own layout.  The fused sweeps
   of the time step solve the same systems with the partition/PCR solver
   instead; these serve callers of the line routines and tests. */

/* CTRIDIAG(A,B,NN), tridiag.f:43-58: A(nsys,NN,3) real (overwritten as
   there), B(nsys,NN) complex (re,im interleaved), solved in place */
int vk_ctridiag_batched(double *a, double *b, const int *nn, const int *nsys);
/* TDMA(a,b,c,r,u,n), periodic_implicit.f:47-64: rows 0..n+1 */
int vk_tdma_batched(const double *a, const double *b, const double *c, const double *r,
                    double *u, const int *n, const int *nsys);
/* cyclic(a,b,c,alpha,beta,r,x,n), periodic_implicit.f:67-93 */
int vk_cyclic_batched(const double *a, const double *b, const double *c, const double *alpha,
                      const double *beta, const double *r, double *x, const int *n,
                      const int *nsys);
/* implicit_solver(SUX,nx,dx,dx12,Re,dt,ft,fb,eta,CUX), periodic_implicit.f:1-44:
   SUX(nsys,0:nx+1) in/out, CUX(nsys,1:nx) */
int vk_implicit_solver_batched(double *sux, const int *nx, const double *dx,
                               const double *dx12, const double *re, const double *dt,
                               const double *ft, const double *fb, const double *eta,
                               const double *cux, const int *nsys);
/* implicit_uy(S,ny,dy,dy12,Re,dt,ft,fb,eta,C), nse_dominik.f:1505-1577:
   S(nsys,0:ny+1) in/out, C(nsys,1:ny) */
int vk_implicit_uy_batched(double *s, const int *ny, const double *dy, const double *dy12,
                           const double *re, const double *dt, const double *ft,
                           const double *fb, const double *eta, const double *c,
                           const int *nsys);
/* implicit_vy(S,ny,dy,dy12,Re,dt,ft,fb,eta,C), nse_dominik.f:1583-1648:
   S(nsys,0:ny) in/out, C(nsys,1:ny-1) */
int vk_implicit_vy_batched(double *s, const int *ny, const double *dy, const double *dy12,
                           const double *re, const double *dt, const double *ft,
                           const double *fb, const double *eta, const double *c,
                           const int *nsys);
/* implicit_Tx(STX,nx,dx,dx12,Re,Sc,dt,llx,rrx,ft,fb,eta,CUX),
   nse_dominik.f:1725-1781: STX, llx, rrx, CUX all (nsys,0:nx+1) */
int vk_implicit_tx_batched(double *stx, const int *nx, const double *dx, const double *dx12,
                           const double *re, const double *sc, const double *dt,
                           const double *llx, const double *rrx, const double *ft,
                           const double *fb, const double *eta, const double *cux,
                           const int *nsys);
/* implicit_Ty(STY,ny,dy,dy12,Re,Sc,dt,lly,rry,ft,fb,CUY,eta),
   nse_dominik.f:1784-1870: STY, lly, rry, CUY all (nsys,0:ny+1); rows
   1..ny of STY are written */
int vk_implicit_ty_batched(double *sty, const int *ny, const double *dy, const double *dy12,
                           const double *re, const double *sc, const double *dt,
                           const double *lly, const double *rry, const double *ft,
                           const double *fb, const double *cuy, const double *eta,
                           const int *nsys);

/* -------------------------------------------- iterative Poisson solver
   The reference's alternative to MKP's spectral solve, which it keeps but
   does not call.  Both keep the reference's arithmetic bit for bit:
   line Gauss-Seidel in the reference's loop order (lines of one
   anti-diagonal of the two outer loop indices solved together), each line
   with its last row's metrics on every row, zero-gradient halos on all six
   faces, p shifted so that p(1,1,1) = 0.  The reference declares its line
   buffers fpy, fpz with nx+2 elements (undefined for ny or nz above nx+1);
   here they hold whole lines. */

/* linebyline(nx,ny,nz,u_temp,v_temp,w_temp,dx,dy,dz,dx12,dy12,dz12,dt,p),
   nse_dominik.f:2136-2267 with implicit_px/py/pz (:2270-2457): one
   iteration on p for L p = div(u_temp)/dt */
int vk_linebyline(const int *nx, const int *ny, const int *nz, const double *u_temp,
                  const double *v_temp, const double *w_temp, const double *dx, const double *dy,
                  const double *dz, const double *dx12, const double *dy12, const double *dz12,
                  const double *dt, double *p);
/* solve_p(p,u_temp,v_temp,w_temp,dt,dx,dy,dz,dx12,dy12,dz12,nx,ny,nz,res,
   res_max,res_int,res_var), nse_dominik.f:1955-2021, plus two trailing
   arguments: linebyline until the largest residual |L p dt - div u_temp|
   is below 1e-5 or max_iterations iterations (NULL: the reference's
   100000); iterations (may be NULL) returns the count.  res is the last
   interior point's residual, res_int and res_var the volume integrals of
   the residual and its square (summed in a fixed tree order: equal to the
   reference's serial sums to rounding); res_max is exact. */
int vk_solve_p(double *p, const double *u_temp, const double *v_temp, const double *w_temp,
               const double *dt, const double *dx, const double *dy, const double *dz,
               const double *dx12, const double *dy12, const double *dz12, const int *nx,
               const int *ny, const int *nz, double *res, double *res_max, double *res_int,
               double *res_var, const int *max_iterations, int *iterations);

/* ------------------------------------------------- pressure projection */
/* MKP, mkp.f:9-220 (ompnthreads is accepted and ignored, as there) */
int vk_mkp(double *p, const double *u_temp, const double *v_temp,
           const double *w_temp, const double *dt, const double *dx,
           const double *dy, const double *dz, const double *dx12,
           const double *dy12, const double *dz12, const double *re,
           const int *nx, const int *ny, const int *nz,
           const int *ompnthreads, double *u_ta, double *v_ta, double *w_ta,
           double *u_ep2, double *v_ep2, double *w_ep2, double *divu,
           double *u, double *v, double *w, const double *xl,
           const double *zl);
/* P2S(FP,FPT,NX,NY,NZ), p2s_fftw2d.f:5-114 */
int vk_p2s(const double *fp, double *fpt, const int *nx, const int *ny,
           const int *nz);
/* S2P(FPT,FP,NX,NY,NZ), p2s_fftw2d.f:116-223 */
int vk_s2p(const double *fpt, double *fp, const int *nx, const int *ny,
           const int *nz);
/* SPN(NX,NY,NZ,AC,KXZ,PHIT), mkp.f:223-260; AC(1:ny,3), KXZ(0:nx/2,0:nz) */
int vk_spn(const int *nx, const int *ny, const int *nz, const double *ac,
           const double *kxz, double *phit);
/* SP0B(NX,NY,NZ,AC,PHIT), mkp.f:263-290 */
int vk_sp0b(const int *nx, const int *ny, const int *nz, const double *ac,
            double *phit);
/* PHIBC(NX,NY,NZ,PHI), mkp.f:293-320 */
int vk_phibc(const int *nx, const int *ny, const int *nz, double *phi);

/* ---------------------------------------------------------- level set */
/* set_levelset, set_levelset.f:1-199 (2001 reinitialisation sweeps; the
   VTK dumps and the printed residual are host I/O and not performed) */
int vk_set_levelset(const int *nx, const int *ny, const int *nz,
                    const double *dx, const double *dx12, const double *dy12,
                    const double *dz12, const double *phi0, double *phi,
                    const double *x, const double *y, const double *z,
                    double *cellvalue);

/* The analytic geometry front-end: phi0 of set_levelset from a union of
   ellipsoids, on the device.  No reference counterpart: the reference
   reads phi0 from structure.dat (READGRID3D, set_levelset.f:202-214), and
   this writes the same array (interior; the halo is set to 0, as a
   zeroed phi0 read by READGRID3D has it).  shapes is host memory,
   nshapes records of 10 doubles: centre (cx, cy, cz), semi-axes (a, b, c)
   along the body axes, and the body rotation as a quaternion (w, x, y, z)
   (normalised here; (1,0,0,0) for axis-aligned).  At each cell centre
   (x(i), y(j), z(k)),
     phi0 = max over shapes of min(a,b,c) (1 - |R^T d / (a,b,c)|),
   d the offset from the centre with the minimum image in the periodic x
   (period xl) and z (period zl): positive inside (the solid), exactly the
   signed distance for a sphere, and -1 everywhere with no shape.  The
   reinitialisation of set_levelset then makes it a signed distance. */
int vk_levelset_shapes(const int *nx, const int *ny, const int *nz, const double *x,
                       const double *y, const double *z, const double *xl, const double *zl,
                       const int *nshapes, const double *shapes, double *phi0);

/* ------------------------------------------------------ adjoint sweeps */
/* set_adj_convect, adjoint.f:1-126 */
int vk_set_adj_convect(const double *u, double *u_v, const double *v,
                       double *v_v, const double *w, double *w_v,
                       const double *re, const double *dt, const int *nx,
                       const int *ny, const int *nz, const double *dx,
                       const double *dy, const double *dz, const double *dx12,
                       const double *dy12, const double *dz12,
                       const double *u_final, const double *v_final,
                       const double *w_final);
/* set_adj_convection_for_scalar, adjoint.f:128-181 */
int vk_set_adj_convection_for_scalar(
    const double *u, const double *v, const double *w, const double *re,
    const double *dt, const int *nx, const int *ny, const int *nz,
    const double *dx, const double *dy, const double *dz, const double *dx12,
    const double *dy12, const double *dz12, const double *t, double *u_vt,
    const double *rhocp, const double *u_final, const double *v_final,
    const double *w_final);
/* set_adj_NS_scalar, adjoint.f:185-211, plus a trailing T_final argument
   (the reference reads an uninitialised local of that name) */
int vk_set_adj_ns_scalar(const double *dt, const int *nx, const int *ny,
                         const int *nz, const double *dx12,
                         const double *dy12, const double *dz12,
                         const double *t, double *u_s, double *v_s,
                         double *w_s, const double *u_final,
                         const double *v_final, const double *w_final,
                         const double *t_final);

/* ------------------------------------------------------ the time step */
/* One pass of the main-loop body nse_dominik.f:290-448 (SET_BC through
   MKP) with the phases fused: no reference routine of this name exists;
   it is what program main's loop does per step.  Reads and writes the
   same arrays the loop does.  u_ep2/v_ep2/w_ep2 (MKP's diagnostic
   splitting error) are skipped when passed as NULL. */
int vk_time_step(double *u, double *v, double *w, double *p, double *t,
                 double *u_pre, double *v_pre, double *w_pre, double *t_pre,
                 double *u_ta, double *v_ta, double *w_ta, double *u_temp,
                 double *v_temp, double *w_temp, double *divu,
                 double *cellvalue, const double *rhocp, const double *thc,
                 double *u_ep2, double *v_ep2, double *w_ep2,
                 const double *dx, const double *dy, const double *dz,
                 const double *dx12, const double *dy12, const double *dz12,
                 const int *nx, const int *ny, const int *nz,
                 const double *dt, const double *re, const double *sc,
                 const double *eta, const double *xl, const double *zl,
                 const int *l);

/* The mixed-precision form of vk_time_step (no reference counterpart:
   the reference is FP64 throughout).  Same arrays and layouts, with these
   held in float: t, u_pre, v_pre, w_pre, t_pre, u_ta, v_ta, w_ta,
   cellvalue, rhocp, thc.  u, v, w, p, u_temp, v_temp, w_temp and divu stay
   double.  The explicit phase and the FFTs compute in FP32, the line
   solves in FP64 on the mixed storage; the projection adds one FP64
   iterative-refinement pass, so the post-projection divergence is at
   FP64 round-off.  When the divergence after that pass is still above
   9e-11 (steps where the penalisation cuts the velocity sharply at the
   solid's surface, such as the first steps from a non-solenoidal state),
   one more refinement pass on the post-projection divergence follows and
   the projection is redone; deciding this reads one word back to the
   host, so the call returns after the projection has run.  MKP's u_ep2
   diagnostic is not produced. */
int vk_time_step_mixed(double *u, double *v, double *w, double *p, float *t,
                       float *u_pre, float *v_pre, float *w_pre, float *t_pre,
                       float *u_ta, float *v_ta, float *w_ta, double *u_temp,
                       double *v_temp, double *w_temp, double *divu,
                       float *cellvalue, const float *rhocp, const float *thc,
                       const double *dx, const double *dy, const double *dz,
                       const double *dx12, const double *dy12,
                       const double *dz12, const int *nx, const int *ny,
                       const int *nz, const double *dt, const double *re,
                       const double *sc, const double *eta, const double *xl,
                       const double *zl, const int *l);

/* ------------------------------------------------- adjoint gradient */
/* No reference counterpart: adjoint.f holds only operator kernels (above),
   with no driver, objective or gradient.

   The objective is the wall heat flux through both walls, integrated over
   the walls and over the run:
     J = sum_{n=1..nsteps} dt sum_{i,k} dx(i) dz(k)
           [2 T^n(i,1,k)/dy12(0) + 2 T^n(i,ny,k)/dy12(ny)],
   T^n the temperature after step n of the FP64 fused step (vk_time_step)
   from the given initial state (arrays as for vk_time_step; none is
   modified).  cellvalue slot 1 is the control: slots 2-4 are derived from
   it at step 1 as SET_IMMERSED_BOUNDARY does.

   vk_objective returns J.  vk_adjoint_gradient returns J and grad =
   dJ/dcellvalue(:,:,:,1) (nx*ny*nz doubles, x fastest): the exact
   discrete adjoint of the forward step, run backward over the trajectory
   with a checkpoint every `checkpoint_every` steps held in device memory
   and each segment recomputed from its checkpoint.  Where a segment's
   tapes fit in device memory (23 fields per step, with 4 GiB to spare),
   every step of the segment is taped as it is recomputed; otherwise each
   step is rerun in the backward pass with one step's tape.  The two give
   the same gradient bit for bit; VK_ADJOINT_SEGMENT_TAPES=0 in the
   environment asks for the second. */
int vk_objective(const double *u, const double *v, const double *w,
                 const double *p, const double *t, const double *u_pre,
                 const double *v_pre, const double *w_pre, const double *t_pre,
                 const double *u_ta, const double *v_ta, const double *w_ta,
                 const double *cellvalue, const double *rhocp,
                 const double *thc, const double *dx, const double *dy,
                 const double *dz, const double *dx12, const double *dy12,
                 const double *dz12, const int *nx, const int *ny,
                 const int *nz, const double *dt, const double *re,
                 const double *sc, const double *eta, const double *xl,
                 const double *zl, const int *nsteps, double *objective);
int vk_adjoint_gradient(const double *u, const double *v, const double *w,
                        const double *p, const double *t, const double *u_pre,
                        const double *v_pre, const double *w_pre,
                        const double *t_pre, const double *u_ta,
                        const double *v_ta, const double *w_ta,
                        const double *cellvalue, const double *rhocp,
                        const double *thc, const double *dx, const double *dy,
                        const double *dz, const double *dx12,
                        const double *dy12, const double *dz12, const int *nx,
                        const int *ny, const int *nz, const double *dt,
                        const double *re, const double *sc, const double *eta,
                        const double *xl, const double *zl, const int *nsteps,
                        const int *checkpoint_every, double *objective,
                        double *grad);

/* The same gradient over a trajectory kept by binomial checkpointing, and
   optionally about the mixed-precision trajectory.  No reference
   counterpart (see vk_adjoint_gradient).
   - snapshots: state slots besides the initial state (which is read from
     the caller's arrays); a negative value takes as many as the free device
     memory holds, keeping 4 GiB.  The schedule is the binomial one
     (Griewank-Walther): the positions minimising the forward steps
     recomputed, found exactly by dynamic programming, so nsteps steps
     reverse in memory for `snapshots` states.  With mixed = 0 the
     gradient is vk_adjoint_gradient's, bit for bit.
   - mixed != 0: the forward trajectory is vk_time_step_mixed's (T, the
     history, u_ta, cellvalue, rhoCp, thc in float; snapshots 3 double
     and 8 float fields); J is that trajectory's, and the backward pass is
     the FP64 adjoint linearised about it.
   - stats (may be NULL): 4 doubles -- snapshots used, forward steps run
     (the objective pass included), the schedule's recomputed steps, device
     bytes the call held. */
int vk_adjoint_gradient_revolve(
    const double *u, const double *v, const double *w, const double *p, const double *t,
    const double *u_pre, const double *v_pre, const double *w_pre, const double *t_pre,
    const double *u_ta, const double *v_ta, const double *w_ta, const double *cellvalue,
    const double *rhocp, const double *thc, const double *dx, const double *dy, const double *dz,
    const double *dx12, const double *dy12, const double *dz12, const int *nx, const int *ny,
    const int *nz, const double *dt, const double *re, const double *sc, const double *eta,
    const double *xl, const double *zl, const int *nsteps, const int *snapshots,
    const int *mixed, double *objective, double *grad, double *stats);

/* dst := src over a whole halo field, or over its interior when
   *interior_only is non-zero: the inline array copies of program main
   (u_pre := u at nse_dominik.f:246-257, u_temp := u_a at :407-417,
   T := u_T at :429-437), which a host with device-resident arrays cannot
   do itself.  No reference routine of this name. */
int vk_field_copy(const double *src, double *dst, const int *nx,
                  const int *ny, const int *nz, const int *interior_only);

/* Max |x| over the interior of a halo field and a NaN/Inf flag
   (returns VK_ERR_NONFINITE when one is found).  No reference
   counterpart: a device reduction for run checks and the divergence
   residual. */
int vk_field_absmax(const double *f, const int *nx, const int *ny,
                    const int *nz, double *result);

/* ------------------------------------------------------- diagnostics
   The reference's run-time checks as device reductions over resident
   fields (metrics resident too); the scalars are host variables, as in the
   reference's calls.  Maxima and positions are exact; sums agree with the
   reference's serial loops to rounding. */

/* CHECK_COURANT(DT,DX,DY,DZ,NX,NY,NZ,U,V,W) (nse_dominik.f:2519-2601,
   called at :467).  The reference prints the largest signed u dt/dx,
   v dt/dy, w dt/dz (from 0); here it is returned in *courant. */
int vk_check_courant(const double *dt, const double *dx, const double *dy,
                     const double *dz, const int *nx, const int *ny,
                     const int *nz, const double *u, const double *v,
                     const double *w, double *courant);

/* vk_check_courant that allocates nothing.  No counterpart in the
   reference: it is for callers that cannot allocate device memory (code
   under HIP graph capture, tight serving loops, harnesses that forbid
   hipMalloc).  Same arguments, kernels and result as vk_check_courant
   (bit for bit), with the per-block partial maxima in a static device
   array of the library instead of workspace: 8 + 3 x 1024 doubles
   (24.6 KB) on each device, held until the process ends.  The pass's
   grid is capped at 1024 blocks, so that serves every grid the
   diagnostics take (fewer than 2^53 interior points); a larger need would
   be refused with VK_ERR_BAD_SIZE.  Every call on a device shares that
   device's array and nothing in it needs clearing between calls; calls
   that overlap on one device (from several host threads) race.  The
   device current on the calling thread must be the library's.  Like
   vk_check_courant it waits for the result. */
int vk_check_courant_noalloc(const double *dt, const double *dx, const double *dy,
                             const double *dz, const int *nx, const int *ny,
                             const int *nz, const double *u, const double *v,
                             const double *w, double *courant);

/* MONITOR(...) (nse_dominik.f:2797-2872, called at :493), same arguments
   and updates: the velocity integrals and their relative changes, T_all
   over the plane k = nz+1 (the reference's loop index after its k loop),
   dif, T_b, and the largest |T - T_before|.  With *int_report == 1 it
   prints what the reference prints.  loc (may be null) receives the
   position the reference prints as xx, yy, zz (0s when T == T_before). */
int vk_monitor(const int *nx, const int *ny, const int *nz, const int *l,
               const double *t, const double *dx, const double *dy,
               const double *dz, double *dif, double *t_all, double *t_dif,
               const double *t_before, const double *u, const double *v,
               const double *w, double *vel_dif, double *u_dif, double *v_dif,
               double *w_dif, double *u_int_before, double *v_int_before,
               double *w_int_before, double *t_b, const double *dt,
               const int *int_report, int *loc);

/* CHECK_CONVERGENCE(Y,YL,U,L,DT,RE,NX,NY,NZ) (nse_dominik.f:2875-2892,
   called at :483): the sum over the plane i = nx of |u - (-Re/2 y (y-yl))|
   that the reference appends to timely_convergence.dat with l*dt; here it
   is returned in *error_sum and the caller writes the record. */
int vk_check_convergence(const double *y, const double *yl, const double *u,
                         const int *l, const double *dt, const double *re,
                         const int *nx, const int *ny, const int *nz,
                         double *error_sum);

/* READGRID3D(NX,NY,NZ,PHI0) (set_levelset.f:202-214, called at
   nse_dominik.f:203): the interior of phi0 from the list-directed records
   "a b c phi0" of structure.dat in the working directory, k, j, i order.
   phi0 is host memory as in the reference; a registered array's device
   mirror is refreshed.  Repeat counts, null values and '/' are refused
   (VK_ERR_ARGUMENT), as are a missing or short file.  Needs no GPU for an
   unregistered array. */
int vk_readgrid3d(const int *nx, const int *ny, const int *nz, double *phi0);

/* OUTFIELD(X,Y,Z,U,V,W,P,NX,NY,NZ,TID,T,PHI,U_TA,V_TA,W_TA)
   (nse_dominik.f:2726-2766, called at :489 and :523): field.u (the record
   tid, then one record of u, v, w, p, T, phi, u_ta, v_ta, w_ta per halo
   point, k, j, i order) and ./nseoutput2/grid.u (x, y, z of the interior,
   one record each), Fortran sequential unformatted with 4-byte record
   markers, byte for byte as the reference writes them.  The fields are
   host arrays; registered ones are downloaded first. */
int vk_outfield(const double *x, const double *y, const double *z,
                const double *u, const double *v, const double *w,
                const double *p, const int *nx, const int *ny, const int *nz,
                const double *tid, const double *t, const double *phi,
                const double *u_ta, const double *v_ta, const double *w_ta);

/* READFIELD(U,V,W,P,NX,NY,NZ,T,PHI,U_TA,V_TA,W_TA) (nse_dominik.f:2769-2793,
   the restart at :240): ./nseoutput2/field.u, one record of 9 values per
   halo point, as OUTPUT writes it.  Unlike the reference's read, a leading
   8-byte record (the tid OUTFIELD writes first) is skipped, so OUTFIELD's
   field.u reads back too.
   Registered arrays are uploaded after the read. */
int vk_readfield(double *u, double *v, double *w, double *p, const int *nx,
                 const int *ny, const int *nz, double *t, double *phi,
                 double *u_ta, double *v_ta, double *w_ta);

/* OUTPUT(U,V,W,P,X,Y,Z,NX,NY,NZ,T,CELLVALUE,U_TA,V_TA,W_TA,PHI,DIVU)
   (nse_dominik.f:2606-2723, called at :480 and :527): creates nseoutput2/
   when missing and writes its twelve unformatted per-point files, gridnumber.dat,
   field.u (no tid record: the file READFIELD reads) and grid.u, byte for
   byte as the reference; temperature_check.dat holds the same numbers in
   list-directed text (17 significant digits).  Registered arrays are
   downloaded first. */
int vk_output(const double *u, const double *v, const double *w,
              const double *p, const double *x, const double *y,
              const double *z, const int *nx, const int *ny, const int *nz,
              const double *t, const double *cellvalue, const double *u_ta,
              const double *v_ta, const double *w_ta, const double *phi,
              const double *divu);

/* The legacy VTK files of the reference, byte for byte: levelset.vtk
   (set_levelset.f:131-156: phi's interior, 'SCALARS phi float64'),
   cellvalue1.vtk (set_levelset.f:171-198) and cellvalue.vtk
   (nse_dominik.f:208-235: cellvalue(...,1), 'SCALARS PHI float').  ASCII
   structured grids with the reference's Fortran edit descriptors:
   'DIMENSIONS ',3(1x,i3) -- which prints *** for an axis of 1000 points or
   more, as the reference does -- 'POINTS ',i9, the cell centres x(i),
   y(j), z(k) as 3e15.7 and the values as e15.7, k, j, i order.  path is a
   NUL-terminated file name (NULL: levelset.vtk / cellvalue.vtk).  Arrays
   are host arrays (registered ones are downloaded first) or device
   memory.  No reference routine of these names: the reference writes the
   files inline. */
int vk_vtk_levelset(const int *nx, const int *ny, const int *nz, const double *x,
                    const double *y, const double *z, const double *phi, const char *path);
int vk_vtk_cellvalue(const int *nx, const int *ny, const int *nz, const double *x,
                     const double *y, const double *z, const double *cellvalue,
                     const char *path);

/* A restart that continues a run bit for bit.  No reference counterpart:
   READFIELD (nse_dominik.f:2769-2793) restores u, v, w, p, T, phi, u_ta,
   v_ta, w_ta but not the AB2 history u_pre, v_pre, w_pre, T_pre nor the
   step counter l, so a run restarted from it takes an Euler step where
   the uninterrupted run takes an AB2 step.  The record written here holds
   everything vk_time_step reads: a header (magic 'VKRESTART 1', nx, ny,
   nz, l), then u, v, w, p, T, u_pre, v_pre, w_pre, T_pre, u_ta, v_ta,
   w_ta, rhoCp, thc (one halo field per record) and cellvalue's four
   slots (one record each), Fortran sequential unformatted with 4-byte
   markers (so each record stays below 2 GiB: grids up to 768x385x768 fit).
   Arrays are host arrays (registered ones are downloaded first / uploaded
   after) or device memory.  vk_read_restart refuses a record for another
   grid or a short file and returns l; the next step is l + 1. */
int vk_write_restart(const char *path, const int *nx, const int *ny, const int *nz,
                     const int *l, const double *u, const double *v, const double *w,
                     const double *p, const double *t, const double *u_pre,
                     const double *v_pre, const double *w_pre, const double *t_pre,
                     const double *u_ta, const double *v_ta, const double *w_ta,
                     const double *rhocp, const double *thc, const double *cellvalue);
int vk_read_restart(const char *path, const int *nx, const int *ny, const int *nz, int *l,
                    double *u, double *v, double *w, double *p, double *t, double *u_pre,
                    double *v_pre, double *w_pre, double *t_pre, double *u_ta, double *v_ta,
                    double *w_ta, double *rhocp, double *thc, double *cellvalue);

#ifdef __cplusplus
}
#endif
#endif
