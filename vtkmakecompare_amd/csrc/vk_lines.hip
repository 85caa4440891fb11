// vk_lines.hip -- the reference's line-level solvers as batched entry
// points: CTRIDIAG (tridiag.f:43-58), TDMA and cyclic (periodic_implicit.f
// :47-93), implicit_solver (periodic_implicit.f:1-44), implicit_uy, _vy,
// _Tx and _Ty (nse_dominik.f:1505-1648, 1725-1870).
//
// The reference solves one line per call.  Here a call solves nsys lines,
// one lane per line, each with the reference's elimination in the
// reference's operation order (the dense (n+2)^2 storage of implicit_uy/vy/Ty
// replaced by its three bands, which are all it touches), so a line's
// result is the reference's bit for bit.  Every array is a Fortran array
// X(nsys, 0:n+1): the systems fastest, so the lanes of a wave read
// consecutive addresses.  The fused sweeps of the time step solve the same
// systems with the partition/PCR solver instead (vk_partition.h); these
// entry points are the reference's own routines for callers and tests.
#include <algorithm>
#include <string>
#include <type_traits>

#include "vk_kernels.h"
#include "vk_runtime.h"

namespace vk {

// element j of line q in an array X(nsys, 0:..)
struct Lane {
  long long q, n;
  __device__ long long operator()(int j) const { return q + n * (long long)j; }
};

// (d1+d2)/(d1/k1+d2/k2): the harmonic face conductance of implicit_Tx/Ty
__device__ inline double hmean_ref(double d1, double k1, double d2, double k2) {
  return (d1 + d2) / (d1 / k1 + d2 / k2);
}

// The band elimination of implicit_uy/vy/Ty (upper sweep from row m down,
// then lower sweep up), rows 0..m, in place on lo/di/up/u
__device__ inline void ul_eliminate(double* lo, double* di, double* up, double* u, Lane L,
                                    int m) {
  for (int j = m; j >= 1; j--) {
    const double aa = di[L(j)];
    di[L(j)] = di[L(j)] / aa;
    lo[L(j)] = lo[L(j)] / aa;
    const double bb = up[L(j - 1)];
    up[L(j - 1)] = up[L(j - 1)] - bb * di[L(j)];
    di[L(j - 1)] = di[L(j - 1)] - bb * lo[L(j)];
    u[L(j)] = u[L(j)] / aa;
    u[L(j - 1)] = u[L(j - 1)] - bb * u[L(j)];
  }
  for (int j = 0; j <= m - 1; j++) {
    const double aa = di[L(j)];
    di[L(j)] = di[L(j)] / aa;
    const double bb = lo[L(j + 1)];
    lo[L(j + 1)] = lo[L(j + 1)] - bb * di[L(j)];
    u[L(j)] = u[L(j)] / aa;
    u[L(j + 1)] = u[L(j + 1)] - bb * u[L(j)];
  }
}

// TDMA (periodic_implicit.f:47-64) on rows 0..n+1; gam is scratch
__device__ inline void tdma(const double* a, const double* b, const double* c, const double* r,
                            double* u, double* gam, Lane L, int n) {
  double bet = b[L(0)];
  u[L(0)] = r[L(0)] / bet;
  for (int j = 1; j <= n + 1; j++) {
    gam[L(j)] = c[L(j - 1)] / bet;
    bet = b[L(j)] - a[L(j)] * gam[L(j)];
    u[L(j)] = (r[L(j)] - a[L(j)] * u[L(j - 1)]) / bet;
  }
  for (int j = n; j >= 0; j--) u[L(j)] = u[L(j)] - gam[L(j + 1)] * u[L(j + 1)];
}

// cyclic (periodic_implicit.f:67-93), Sherman-Morrison with gamma = 1;
// w holds 4 scratch lines (bb, u, z, gam)
__device__ inline void cyclic(const double* a, const double* b, const double* c, double alpha,
                              double beta, const double* r, double* x, double* w, Lane L,
                              int n) {
  double* bb = w;
  double* uu = w + L.n * (long long)(n + 2);
  double* z = uu + L.n * (long long)(n + 2);
  double* gam = z + L.n * (long long)(n + 2);
  const double gamma = 1.0;
  bb[L(0)] = b[L(0)] - gamma;
  bb[L(n + 1)] = b[L(n + 1)] - alpha * beta / gamma;
  for (int i = 1; i <= n; i++) bb[L(i)] = b[L(i)];
  tdma(a, bb, c, r, x, gam, L, n);
  uu[L(0)] = gamma;
  uu[L(n + 1)] = alpha;
  for (int i = 1; i <= n; i++) uu[L(i)] = 0.0;
  tdma(a, bb, c, uu, z, gam, L, n);
  const double fact = (x[L(0)] + beta * x[L(n + 1)] / gamma) /
                      (1.0 + z[L(0)] + beta * z[L(n + 1)] / gamma);
  for (int i = 0; i <= n + 1; i++) x[L(i)] = x[L(i)] - fact * z[L(i)];
}

// the periodic ghost rows of periodic_implicit.f:34-41: x0 = xn, x(n+1) = x1
__device__ inline void periodic_ghost_rows(double* aa, double* bb, double* cc, Lane L, int n) {
  aa[L(0)] = 0.0;
  bb[L(0)] = 0.0;
  cc[L(0)] = 1.0;
  aa[L(n + 1)] = 1.0;
  bb[L(n + 1)] = 0.0;
  cc[L(n + 1)] = 0.0;
}

// CTRIDIAG (tridiag.f:43-58): A(nsys, NN, 3) real, B(nsys, NN) complex; the
// complex/real divisions per component
__global__ void k_ctridiag(double* a, double* b, int nn, int nsys) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nsys) return;
  const Lane L{q, nsys};
  double* a1 = a;
  double* a2 = a + (long long)nsys * nn;
  double* a3 = a + 2 * (long long)nsys * nn;
  for (int i = 0; i < nn - 1; i++) {
    a3[L(i)] = a3[L(i)] / a2[L(i)];
    b[2 * L(i)] = b[2 * L(i)] / a2[L(i)];
    b[2 * L(i) + 1] = b[2 * L(i) + 1] / a2[L(i)];
    a2[L(i + 1)] = a2[L(i + 1)] - a1[L(i + 1)] * a3[L(i)];
    b[2 * L(i + 1)] = b[2 * L(i + 1)] - a1[L(i + 1)] * b[2 * L(i)];
    b[2 * L(i + 1) + 1] = b[2 * L(i + 1) + 1] - a1[L(i + 1)] * b[2 * L(i) + 1];
  }
  b[2 * L(nn - 1)] = b[2 * L(nn - 1)] / a2[L(nn - 1)];
  b[2 * L(nn - 1) + 1] = b[2 * L(nn - 1) + 1] / a2[L(nn - 1)];
  for (int i = nn - 2; i >= 0; i--) {
    b[2 * L(i)] = b[2 * L(i)] - a3[L(i)] * b[2 * L(i + 1)];
    b[2 * L(i) + 1] = b[2 * L(i) + 1] - a3[L(i)] * b[2 * L(i + 1) + 1];
  }
}

__global__ void k_tdma(const double* a, const double* b, const double* c, const double* r,
                       double* u, double* gam, int n, int nsys) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nsys) return;
  tdma(a, b, c, r, u, gam, Lane{q, nsys}, n);
}

__global__ void k_cyclic(const double* a, const double* b, const double* c, const double* alpha,
                         const double* beta, const double* r, double* x, double* w, int n,
                         int nsys) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nsys) return;
  cyclic(a, b, c, alpha[q], beta[q], r, x, w, Lane{q, nsys}, n);
}

// implicit_solver (periodic_implicit.f:1-44): SUX(nsys, 0:nx+1) in/out,
// CUX(nsys, 1:nx); w holds 8 scratch lines
__global__ void k_implicit_solver(double* sux, int nx, const double* dx, const double* dx12,
                                  double re, double dt, const double* ft, const double* fb,
                                  double eta, const double* cux, double* w, int nsys) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nsys) return;
  const Lane L{q, nsys};
  const long long ln = (long long)nsys * (nx + 2);
  double *u = w, *aa = w + ln, *bb = w + 2 * ln, *cc = w + 3 * ln;
  const double beta = dt / re / 2.0;
  for (int i = 1; i <= nx; i++) u[L(i)] = sux[L(i)];
  u[L(0)] = ft[q];
  u[L(nx + 1)] = fb[q];
  for (int i = 1; i <= nx; i++) {
    const double alpha = dt * eta * cux[L(i - 1)] + 1.0;
    aa[L(i)] = -beta / alpha / dx[i] / dx12[i];
    bb[L(i)] = 1.0 + beta / alpha * (1.0 / dx[i + 1] + 1.0 / dx[i]) / dx12[i];
    cc[L(i)] = -beta / alpha / dx[i + 1] / dx12[i];
  }
  periodic_ghost_rows(aa, bb, cc, L, nx);
  cyclic(aa, bb, cc, -1.0, -1.0, u, sux, w + 4 * ln, L, nx);
}

// implicit_uy (nse_dominik.f:1505-1577): S(nsys, 0:ny+1), C(nsys, 1:ny);
// w holds 4 scratch lines
__global__ void k_implicit_uy(double* s, int ny, const double* dy, const double* dy12,
                              double re, double dt, const double* ft, const double* fb,
                              double eta, const double* c, double* w, int nsys) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nsys) return;
  const Lane L{q, nsys};
  const int m = ny + 1;
  const long long ln = (long long)nsys * (m + 1);
  double *lo = w, *di = w + ln, *up = w + 2 * ln, *u = w + 3 * ln;
  const double beta = dt / re / 2.0;
  for (int j = 1; j <= ny; j++) {
    const double alpha = dt * eta * c[L(j - 1)] + 1.0;
    lo[L(j)] = -beta / alpha / dy12[j - 1] / dy[j];
    di[L(j)] = 1.0 + beta / alpha * (1.0 / dy12[j] + 1.0 / dy12[j - 1]) / dy[j];
    up[L(j)] = -beta / alpha / dy12[j] / dy[j];
  }
  lo[L(0)] = 0.0;
  di[L(0)] = 1.0;
  up[L(0)] = 1.0;
  di[L(ny + 1)] = 1.0;
  lo[L(ny + 1)] = 1.0;
  up[L(ny + 1)] = 0.0;
  for (int j = 1; j <= ny; j++) u[L(j)] = s[L(j)];
  u[L(0)] = ft[q];
  u[L(ny + 1)] = fb[q];
  ul_eliminate(lo, di, up, u, L, m);
  for (int j = 0; j <= ny + 1; j++) s[L(j)] = u[L(j)];
}

// implicit_vy (nse_dominik.f:1583-1648): S(nsys, 0:ny), C(nsys, 1:ny-1)
__global__ void k_implicit_vy(double* s, int ny, const double* dy, const double* dy12,
                              double re, double dt, const double* ft, const double* fb,
                              double eta, const double* c, double* w, int nsys) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nsys) return;
  const Lane L{q, nsys};
  const int m = ny;
  const long long ln = (long long)nsys * (m + 1);
  double *lo = w, *di = w + ln, *up = w + 2 * ln, *u = w + 3 * ln;
  const double beta = dt / re / 2.0;
  lo[L(0)] = 0.0;
  di[L(0)] = 1.0;
  up[L(0)] = 0.0;
  lo[L(ny)] = 0.0;
  di[L(ny)] = 1.0;
  up[L(ny)] = 0.0;
  for (int j = 1; j <= ny - 1; j++) {
    const double alpha = dt * eta * c[L(j - 1)] + 1.0;
    lo[L(j)] = -beta / alpha / dy[j] / dy12[j];
    di[L(j)] = 1.0 + beta / alpha * (1.0 / dy[j + 1] + 1.0 / dy[j]) / dy12[j];
    up[L(j)] = -beta / alpha / dy[j + 1] / dy12[j];
  }
  for (int j = 1; j <= ny - 1; j++) u[L(j)] = s[L(j)];
  u[L(0)] = ft[q];
  u[L(ny)] = fb[q];
  ul_eliminate(lo, di, up, u, L, m);
  for (int j = 0; j <= ny; j++) s[L(j)] = u[L(j)];
}

// implicit_Tx (nse_dominik.f:1725-1781): STX, LLX, RRX, CUX all (nsys, 0:nx+1)
__global__ void k_implicit_tx(double* stx, int nx, const double* dx, const double* dx12,
                              double re, double sc, double dt, const double* llx,
                              const double* rrx, const double* ft, const double* fb, double eta,
                              const double* cux, double* w, int nsys) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nsys) return;
  const Lane L{q, nsys};
  const long long ln = (long long)nsys * (nx + 2);
  double *tt = w, *aa = w + ln, *bb = w + 2 * ln, *cc = w + 3 * ln;
  const double beta = dt / re / sc / 2.0;
  for (int i = 1; i <= nx; i++) tt[L(i)] = stx[L(i)];
  tt[L(0)] = ft[q];
  tt[L(nx + 1)] = fb[q];
  for (int i = 1; i <= nx; i++) {
    const double alpha = dt * eta * cux[L(i)] + 1.0;
    const double h1 = hmean_ref(dx[i], llx[L(i)], dx[i - 1], llx[L(i - 1)]);
    const double h2 = hmean_ref(dx[i + 1], llx[L(i + 1)], dx[i], llx[L(i)]);
    aa[L(i)] = -beta / alpha * h1 / dx12[i - 1] / dx[i] / rrx[L(i)];
    bb[L(i)] = 1.0 + beta / alpha * (h2 / dx12[i] / dx[i] + h1 / dx12[i - 1] / dx[i]) /
                         rrx[L(i)];
    cc[L(i)] = -beta / alpha * h2 / dx12[i] / dx[i] / rrx[L(i)];
  }
  periodic_ghost_rows(aa, bb, cc, L, nx);
  cyclic(aa, bb, cc, -1.0, -1.0, tt, stx, w + 4 * ln, L, nx);
}

// implicit_Ty (nse_dominik.f:1784-1870): STY, LLY, RRY, CUY all
// (nsys, 0:ny+1); STY's rows 1..ny are written
__global__ void k_implicit_ty(double* sty, int ny, const double* dy, const double* dy12,
                              double re, double sc, double dt, const double* lly,
                              const double* rry, const double* ft, const double* fb,
                              const double* cuy, double eta, double* w, int nsys) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nsys) return;
  const Lane L{q, nsys};
  const int m = ny + 1;
  const long long ln = (long long)nsys * (m + 1);
  double *lo = w, *di = w + ln, *up = w + 2 * ln, *tt = w + 3 * ln;
  const double beta = dt / re / sc / 2.0;
  lo[L(0)] = 0.0;
  di[L(0)] = 1.0;
  up[L(0)] = 1.0;
  di[L(ny + 1)] = 1.0;
  lo[L(ny + 1)] = 1.0;
  up[L(ny + 1)] = 0.0;
  for (int i = 1; i <= ny; i++) {
    const double alpha = dt * eta * cuy[L(i)] + 1.0;
    const double h1 = hmean_ref(dy[i], lly[L(i)], dy[i - 1], lly[L(i - 1)]);
    const double h2 = hmean_ref(dy[i + 1], lly[L(i + 1)], dy[i], lly[L(i)]);
    lo[L(i)] = -beta / alpha * h1 / dy12[i - 1] / dy[i] / rry[L(i)];
    di[L(i)] = 1.0 + beta / alpha * (h2 / dy12[i] / dy[i] + h1 / dy12[i - 1] / dy[i]) /
                         rry[L(i)];
    up[L(i)] = -beta / alpha * h2 / dy12[i] / dy[i] / rry[L(i)];
  }
  for (int i = 1; i <= ny; i++) tt[L(i)] = sty[L(i)];
  tt[L(0)] = ft[q];
  tt[L(ny + 1)] = fb[q];
  ul_eliminate(lo, di, up, tt, L, m);
  for (int i = 1; i <= ny; i++) sty[L(i)] = tt[L(i)];
}

// ------------------------------------------------ iterative Poisson
// linebyline (nse_dominik.f:2136-2267): a lexicographic line Gauss-Seidel
// sweep.  A line reads its neighbour lines: those before it in the
// reference's loop order already updated, those after it not yet.  Lines
// whose two loop indices sum to the same value read no line of their own
// anti-diagonal, so each anti-diagonal is one launch (a wavefront), and the
// result is the serial sweep's, bit for bit.

// q = div(u_temp)/dt on the interior
__global__ void k_lbl_q(Geom g, const double* u, const double* v, const double* w,
                        const double* dx, const double* dy, const double* dz, double dt,
                        double* q) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  const int k = 1 + blockIdx.z;
  if (i > g.nx || j > g.ny) return;
  const long long o = g.at(i, j, k);
  q[o] = ((u[o] - u[o - 1]) / dx[i] + (v[o] - v[o - g.sy]) / dy[j] +
          (w[o] - w[o - g.sz]) / dz[k]) / dt;
}

struct PMetrics {
  const double *dx, *dy, *dz, *dx12, *dy12, *dz12;
};

// the six coefficients of the Poisson row at (i, j, k)
struct PB {
  double bx1, bx2, by1, by2, bz1, bz2;
  __device__ PB(const PMetrics& m, int i, int j, int k)
      : bx1(1.0 / m.dx12[i] / m.dx[i]), bx2(1.0 / m.dx12[i - 1] / m.dx[i]),
        by1(1.0 / m.dy12[j] / m.dy[j]), by2(1.0 / m.dy12[j - 1] / m.dy[j]),
        bz1(1.0 / m.dz12[k] / m.dz[k]), bz2(1.0 / m.dz12[k - 1] / m.dz[k]) {}
  __device__ double sum() const { return bx1 + bx2 + by1 + by2 + bz1 + bz2; }
};

// One anti-diagonal of the sweep along DIR (0 x lines (j,k), 1 y lines
// (i,k), 2 z lines (i,j)): the line of lane t has loop indices (a, b) =
// (d - b, bmin + t), a the inner loop's.  w: 4 band lines of n+2 values
// per lane, lanes fastest (L lanes).
template <int DIR>
__global__ void k_lbl_sweep(Geom g, PMetrics m, double* p, const double* q, int d, int bmin,
                            int nlines, double* w, int L) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nlines) return;
  const int b = bmin + t, a = d - b;
  const int n = DIR == 0 ? g.nx : DIR == 1 ? g.ny : g.nz;
  const Lane ln{t, L};
  const long long ls = (long long)L * (n + 2);
  double *lo = w, *di = w + ls, *up = w + 2 * ls, *u = w + 3 * ls;
  int i = 0, j = 0, k = 0;
  if (DIR == 0) { j = 1 + a; k = 1 + b; }
  if (DIR == 1) { i = 1 + a; k = 1 + b; }
  if (DIR == 2) { i = 1 + a; j = 1 + b; }
  const long long st = DIR == 0 ? 1 : DIR == 1 ? g.sy : g.sz;
  long long o = g.at(DIR == 0 ? 1 : i, DIR == 1 ? 1 : j, DIR == 2 ? 1 : k);
  for (int e = 1; e <= n; e++, o += st) {
    if (DIR == 0) i = e;
    if (DIR == 1) j = e;
    if (DIR == 2) k = e;
    const PB c(m, i, j, k);
    double f;
    if (DIR == 0)
      f = -c.by1 * p[o + g.sy] - c.by2 * p[o - g.sy] - c.bz1 * p[o + g.sz] - c.bz2 * p[o - g.sz] +
          q[o];
    else if (DIR == 1)
      f = -c.bx1 * p[o + 1] - c.bx2 * p[o - 1] - c.bz1 * p[o + g.sz] - c.bz2 * p[o - g.sz] + q[o];
    else
      f = -c.bx1 * p[o + 1] - c.bx2 * p[o - 1] - c.by1 * p[o + g.sy] - c.by2 * p[o - g.sy] + q[o];
    u[ln(e)] = f;
  }
  // the line's matrix: every row with the coefficients of its last row
  const PB c(m, DIR == 0 ? g.nx : i, DIR == 1 ? g.ny : j, DIR == 2 ? g.nz : k);
  const double blo = DIR == 0 ? c.bx2 : DIR == 1 ? c.by2 : c.bz2;
  const double bhi = DIR == 0 ? c.bx1 : DIR == 1 ? c.by1 : c.bz1;
  const double bsum = c.sum();
  lo[ln(0)] = 0.0;
  di[ln(0)] = 1.0;
  up[ln(0)] = -1.0;
  lo[ln(n + 1)] = -1.0;
  di[ln(n + 1)] = 1.0;
  up[ln(n + 1)] = 0.0;
  for (int e = 1; e <= n; e++) {
    lo[ln(e)] = blo;
    di[ln(e)] = -(bsum);
    up[ln(e)] = bhi;
  }
  u[ln(0)] = 0.0;
  u[ln(n + 1)] = 0.0;
  ul_eliminate(lo, di, up, u, ln, n + 1);
  o = g.at(DIR == 0 ? 1 : i, DIR == 1 ? 1 : j, DIR == 2 ? 1 : k);
  for (int e = 1; e <= n; e++, o += st) p[o] = u[ln(e)];
}

// zero-gradient halos in the reference's order: z faces, x faces, y faces
// (each over the other two indices' full ranges)
template <int FACE>
__global__ void k_lbl_halo(Geom g, double* p) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (FACE == 0) {  // (i, j) over 0..nx+1, 0..ny+1
    if (a > g.nx + 1) return;
    p[g.at(a, b, 0)] = p[g.at(a, b, 1)];
    p[g.at(a, b, g.nz + 1)] = p[g.at(a, b, g.nz)];
  } else if (FACE == 1) {  // (j, k)
    if (a > g.ny + 1) return;
    p[g.at(0, a, b)] = p[g.at(1, a, b)];
    p[g.at(g.nx + 1, a, b)] = p[g.at(g.nx, a, b)];
  } else {  // (i, k)
    if (a > g.nx + 1) return;
    p[g.at(a, 0, b)] = p[g.at(a, 1, b)];
    p[g.at(a, g.ny + 1, b)] = p[g.at(a, g.ny, b)];
  }
}

__global__ void k_lbl_pst(Geom g, const double* p, double* pst) { *pst = p[g.at(1, 1, 1)]; }

__global__ void k_lbl_shift(long long n, double* p, const double* pst) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) p[q] = p[q] - *pst;
}

// solve_p's residual statistics (nse_dominik.f:1982-2001), pass 1: per
// block, over the interior in linear order, the largest |L p dt - div| and
// the sums of (L p dt - div) dV and |.|^2 dV; block 0 also keeps the last
// point's |.|.  Pass 2 combines the blocks in order.
//
// The statistics share one workspace of kLblResWords doubles with the
// mean-pressure word: [0] the pressure at (1,1,1), [4..7] the four
// results, then from kLblPart the per-block triples (3 kResBlocks words)
// and, after them, the last point's residual.
constexpr int kResBlocks = 1024;
constexpr int kLblOut = 4, kLblPart = 8;
constexpr int kLblLast = 3 * kResBlocks;  // offset from kLblPart
constexpr size_t kLblResWords = kLblPart + kLblLast + 8;
static_assert(kLblOut + 4 <= kLblPart && kLblPart + kLblLast < (int)kLblResWords,
              "solve_p's statistics overrun their workspace");

__global__ __launch_bounds__(256) void k_lbl_resid(Geom g, PMetrics m, const double* u,
                                                   const double* v, const double* w,
                                                   const double* p, double dt, double* part) {
  __shared__ double sm[3][4];
  const long long n = (long long)g.interior();
  double rmax = 0.0, rint = 0.0, rvar = 0.0;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (long long)gridDim.x * blockDim.x) {
    const int i = 1 + (int)(q % g.nx);
    const long long r = q / g.nx;
    const int j = 1 + (int)(r % g.ny), k = 1 + (int)(r / g.ny);
    const long long o = g.at(i, j, k);
    const double div = (u[o] - u[o - 1]) / m.dx[i] + (v[o] - v[o - g.sy]) / m.dy[j] +
                       (w[o] - w[o - g.sz]) / m.dz[k];
    const double pphi =
        ((p[o + 1] - p[o]) / m.dx12[i] - (p[o] - p[o - 1]) / m.dx12[i - 1]) / m.dx[i] +
        ((p[o + g.sy] - p[o]) / m.dy12[j] - (p[o] - p[o - g.sy]) / m.dy12[j - 1]) / m.dy[j] +
        ((p[o + g.sz] - p[o]) / m.dz12[k] - (p[o] - p[o - g.sz]) / m.dz12[k - 1]) / m.dz[k];
    const double res = fabs(pphi * dt - div);
    rmax = res >= rmax ? res : rmax;
    rint = rint + (pphi * dt - div) * m.dx[i] * m.dy[j] * m.dz[k];
    rvar = rvar + (res * res) * m.dx[i] * m.dy[j] * m.dz[k];
    if (q == n - 1) part[kLblLast] = res;  // the last point's residual
  }
  for (int off = 32; off > 0; off >>= 1) {
    const double a = __shfl_xor(rmax, off);
    rmax = a > rmax ? a : rmax;
    rint = rint + __shfl_xor(rint, off);
    rvar = rvar + __shfl_xor(rvar, off);
  }
  if (threadIdx.x % 64 == 0) {
    sm[0][threadIdx.x / 64] = rmax;
    sm[1][threadIdx.x / 64] = rint;
    sm[2][threadIdx.x / 64] = rvar;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 1; q < 4; q++) {
      rmax = sm[0][q] > rmax ? sm[0][q] : rmax;
      rint = rint + sm[1][q];
      rvar = rvar + sm[2][q];
    }
    part[3 * blockIdx.x] = rmax;
    part[3 * blockIdx.x + 1] = rint;
    part[3 * blockIdx.x + 2] = rvar;
  }
}

__global__ void k_lbl_resid_final(const double* part, int nb, double* out) {
  if (threadIdx.x != 0) return;
  double rmax = 0.0, rint = 0.0, rvar = 0.0;
  for (int b = 0; b < nb; b++) {
    rmax = part[3 * b] > rmax ? part[3 * b] : rmax;
    rint = rint + part[3 * b + 1];
    rvar = rvar + part[3 * b + 2];
  }
  out[0] = part[kLblLast];
  out[1] = rmax;
  out[2] = rint;
  out[3] = rvar;
}

// one linebyline iteration on the stream
static int linebyline(const Geom& g, const PMetrics& m, const double* u, const double* v,
                      const double* w, double dt, double* p, double* q, double* band, int L,
                      double* pst, hipStream_t s) {
  k_lbl_q<<<interior_grid(g), interior_block(), 0, s>>>(g, u, v, w, m.dx, m.dy, m.dz, dt, q);
  auto sweep = [&](auto dir, int na, int nb) {
    constexpr int D = decltype(dir)::value;
    for (int d = 0; d <= na + nb - 2; d++) {
      const int bmin = d - (na - 1) > 0 ? d - (na - 1) : 0;
      const int bmax = d < nb - 1 ? d : nb - 1;
      const int nl = bmax - bmin + 1;
      k_lbl_sweep<D><<<(unsigned)((nl + 63) / 64), 64, 0, s>>>(g, m, p, q, d, bmin, nl, band, L);
    }
  };
  sweep(std::integral_constant<int, 0>(), g.ny, g.nz);
  sweep(std::integral_constant<int, 1>(), g.nx, g.nz);
  sweep(std::integral_constant<int, 2>(), g.nx, g.ny);
  k_lbl_halo<0><<<dim3((unsigned)((g.nx + 2 + 63) / 64), (unsigned)(g.ny + 2)), 64, 0, s>>>(g, p);
  k_lbl_halo<1><<<dim3((unsigned)((g.ny + 2 + 63) / 64), (unsigned)(g.nz + 2)), 64, 0, s>>>(g, p);
  k_lbl_halo<2><<<dim3((unsigned)((g.nx + 2 + 63) / 64), (unsigned)(g.nz + 2)), 64, 0, s>>>(g, p);
  k_lbl_pst<<<1, 1, 0, s>>>(g, p, pst);
  const long long nc = (long long)g.cells();
  k_lbl_shift<<<(unsigned)((nc + 255) / 256), 256, 0, s>>>(nc, p, pst);
  return hip_check(hipGetLastError(), "linebyline");
}

constexpr int kLineThreads = 64;

inline unsigned line_blocks(int nsys) { return (unsigned)((nsys + kLineThreads - 1) / kLineThreads); }

// scratch of `lines` lines of n+2 values for nsys systems
static double* line_scratch(int lines, int n, int nsys) {
  return static_cast<double*>(
      workspace(kWorkLineBatch, (size_t)lines * (size_t)nsys * (size_t)(n + 2) * sizeof(double)));
}

}  // namespace vk

using namespace vk;

#define LINE_SIZES(n, what)                                                           \
  if (!(n) || !nsys || *(n) < 1 || *nsys < 1)                                         \
    return fail(VK_ERR_BAD_SIZE, std::string(what) + ": n >= 1 and nsys >= 1");       \
  const size_t ns = (size_t)*nsys
#define LINE_ARRAY(var, type, ptr, count) VK_RESOLVE(var, type, ptr, (count) * sizeof(double))

extern "C" {

int vk_ctridiag_batched(double* a, double* b, const int* nn, const int* nsys) {
  VK_REQUIRE_INIT();
  LINE_SIZES(nn, "ctridiag_batched");
  const size_t n = (size_t)*nn;
  LINE_ARRAY(da, double, a, 3 * ns * n);
  LINE_ARRAY(db, double, b, 2 * ns * n);
  k_ctridiag<<<line_blocks(*nsys), kLineThreads, 0, rt().stream>>>(da, db, *nn, *nsys);
  VK_LAUNCH_CHECK("ctridiag_batched");
  return VK_OK;
}

int vk_tdma_batched(const double* a, const double* b, const double* c, const double* r,
                    double* u, const int* n, const int* nsys) {
  VK_REQUIRE_INIT();
  LINE_SIZES(n, "tdma_batched");
  const size_t m = ns * (size_t)(*n + 2);
  LINE_ARRAY(da, const double, a, m);
  LINE_ARRAY(dbb, const double, b, m);
  LINE_ARRAY(dc, const double, c, m);
  LINE_ARRAY(dr, const double, r, m);
  LINE_ARRAY(du, double, u, m);
  double* w = line_scratch(1, *n, *nsys);
  if (!w) return VK_ERR_HIP;
  k_tdma<<<line_blocks(*nsys), kLineThreads, 0, rt().stream>>>(da, dbb, dc, dr, du, w, *n,
                                                                 *nsys);
  VK_LAUNCH_CHECK("tdma_batched");
  return VK_OK;
}

int vk_cyclic_batched(const double* a, const double* b, const double* c, const double* alpha,
                      const double* beta, const double* r, double* x, const int* n,
                      const int* nsys) {
  VK_REQUIRE_INIT();
  LINE_SIZES(n, "cyclic_batched");
  const size_t m = ns * (size_t)(*n + 2);
  LINE_ARRAY(da, const double, a, m);
  LINE_ARRAY(dbb, const double, b, m);
  LINE_ARRAY(dc, const double, c, m);
  LINE_ARRAY(dal, const double, alpha, ns);
  LINE_ARRAY(dbe, const double, beta, ns);
  LINE_ARRAY(dr, const double, r, m);
  LINE_ARRAY(dx, double, x, m);
  double* w = line_scratch(4, *n, *nsys);
  if (!w) return VK_ERR_HIP;
  k_cyclic<<<line_blocks(*nsys), kLineThreads, 0, rt().stream>>>(da, dbb, dc, dal, dbe, dr, dx,
                                                                   w, *n, *nsys);
  VK_LAUNCH_CHECK("cyclic_batched");
  return VK_OK;
}

int vk_implicit_solver_batched(double* sux, const int* nx, const double* dx, const double* dx12,
                               const double* re, const double* dt, const double* ft,
                               const double* fb, const double* eta, const double* cux,
                               const int* nsys) {
  VK_REQUIRE_INIT();
  LINE_SIZES(nx, "implicit_solver_batched");
  const size_t n2 = (size_t)*nx + 2;
  LINE_ARRAY(ds, double, sux, ns * n2);
  LINE_ARRAY(ddx, const double, dx, n2);
  LINE_ARRAY(ddx12, const double, dx12, n2);
  LINE_ARRAY(dft, const double, ft, ns);
  LINE_ARRAY(dfb, const double, fb, ns);
  LINE_ARRAY(dc, const double, cux, ns * (size_t)*nx);
  double* w = line_scratch(8, *nx, *nsys);
  if (!w) return VK_ERR_HIP;
  k_implicit_solver<<<line_blocks(*nsys), kLineThreads, 0, rt().stream>>>(
      ds, *nx, ddx, ddx12, *re, *dt, dft, dfb, *eta, dc, w, *nsys);
  VK_LAUNCH_CHECK("implicit_solver_batched");
  return VK_OK;
}

int vk_implicit_uy_batched(double* s, const int* ny, const double* dy, const double* dy12,
                           const double* re, const double* dt, const double* ft,
                           const double* fb, const double* eta, const double* c,
                           const int* nsys) {
  VK_REQUIRE_INIT();
  LINE_SIZES(ny, "implicit_uy_batched");
  const size_t n2 = (size_t)*ny + 2;
  LINE_ARRAY(ds, double, s, ns * n2);
  LINE_ARRAY(ddy, const double, dy, n2);
  LINE_ARRAY(ddy12, const double, dy12, n2);
  LINE_ARRAY(dft, const double, ft, ns);
  LINE_ARRAY(dfb, const double, fb, ns);
  LINE_ARRAY(dc, const double, c, ns * (size_t)*ny);
  double* w = line_scratch(4, *ny, *nsys);
  if (!w) return VK_ERR_HIP;
  k_implicit_uy<<<line_blocks(*nsys), kLineThreads, 0, rt().stream>>>(
      ds, *ny, ddy, ddy12, *re, *dt, dft, dfb, *eta, dc, w, *nsys);
  VK_LAUNCH_CHECK("implicit_uy_batched");
  return VK_OK;
}

int vk_implicit_vy_batched(double* s, const int* ny, const double* dy, const double* dy12,
                           const double* re, const double* dt, const double* ft,
                           const double* fb, const double* eta, const double* c,
                           const int* nsys) {
  VK_REQUIRE_INIT();
  LINE_SIZES(ny, "implicit_vy_batched");
  if (*ny < 2) return fail(VK_ERR_BAD_SIZE, "implicit_vy_batched: ny >= 2");
  const size_t n2 = (size_t)*ny + 2;
  LINE_ARRAY(ds, double, s, ns * (n2 - 1));
  LINE_ARRAY(ddy, const double, dy, n2);
  LINE_ARRAY(ddy12, const double, dy12, n2);
  LINE_ARRAY(dft, const double, ft, ns);
  LINE_ARRAY(dfb, const double, fb, ns);
  LINE_ARRAY(dc, const double, c, ns * (size_t)(*ny - 1));
  double* w = line_scratch(4, *ny - 1, *nsys);
  if (!w) return VK_ERR_HIP;
  k_implicit_vy<<<line_blocks(*nsys), kLineThreads, 0, rt().stream>>>(
      ds, *ny, ddy, ddy12, *re, *dt, dft, dfb, *eta, dc, w, *nsys);
  VK_LAUNCH_CHECK("implicit_vy_batched");
  return VK_OK;
}

int vk_implicit_tx_batched(double* stx, const int* nx, const double* dx, const double* dx12,
                           const double* re, const double* sc, const double* dt,
                           const double* llx, const double* rrx, const double* ft,
                           const double* fb, const double* eta, const double* cux,
                           const int* nsys) {
  VK_REQUIRE_INIT();
  LINE_SIZES(nx, "implicit_tx_batched");
  const size_t n2 = (size_t)*nx + 2;
  LINE_ARRAY(ds, double, stx, ns * n2);
  LINE_ARRAY(ddx, const double, dx, n2);
  LINE_ARRAY(ddx12, const double, dx12, n2);
  LINE_ARRAY(dl, const double, llx, ns * n2);
  LINE_ARRAY(drr, const double, rrx, ns * n2);
  LINE_ARRAY(dft, const double, ft, ns);
  LINE_ARRAY(dfb, const double, fb, ns);
  LINE_ARRAY(dc, const double, cux, ns * n2);
  double* w = line_scratch(8, *nx, *nsys);
  if (!w) return VK_ERR_HIP;
  k_implicit_tx<<<line_blocks(*nsys), kLineThreads, 0, rt().stream>>>(
      ds, *nx, ddx, ddx12, *re, *sc, *dt, dl, drr, dft, dfb, *eta, dc, w, *nsys);
  VK_LAUNCH_CHECK("implicit_tx_batched");
  return VK_OK;
}

int vk_implicit_ty_batched(double* sty, const int* ny, const double* dy, const double* dy12,
                           const double* re, const double* sc, const double* dt,
                           const double* lly, const double* rry, const double* ft,
                           const double* fb, const double* cuy, const double* eta,
                           const int* nsys) {
  VK_REQUIRE_INIT();
  LINE_SIZES(ny, "implicit_ty_batched");
  const size_t n2 = (size_t)*ny + 2;
  LINE_ARRAY(ds, double, sty, ns * n2);
  LINE_ARRAY(ddy, const double, dy, n2);
  LINE_ARRAY(ddy12, const double, dy12, n2);
  LINE_ARRAY(dl, const double, lly, ns * n2);
  LINE_ARRAY(drr, const double, rry, ns * n2);
  LINE_ARRAY(dft, const double, ft, ns);
  LINE_ARRAY(dfb, const double, fb, ns);
  LINE_ARRAY(dc, const double, cuy, ns * n2);
  double* w = line_scratch(4, *ny, *nsys);
  if (!w) return VK_ERR_HIP;
  k_implicit_ty<<<line_blocks(*nsys), kLineThreads, 0, rt().stream>>>(
      ds, *ny, ddy, ddy12, *re, *sc, *dt, dl, drr, dft, dfb, dc, *eta, w, *nsys);
  VK_LAUNCH_CHECK("implicit_ty_batched");
  return VK_OK;
}

// the linebyline workspaces: q, the band scratch of the longest
// anti-diagonal, p(1,1,1) and the residual partials
struct LblWork {
  double *q, *band, *pst;
  int L;
};

static int lbl_work(const Geom& g, LblWork& lw) {
  const int nmax = std::max(g.nx, std::max(g.ny, g.nz));
  lw.L = std::max(std::min(g.ny, g.nz), std::max(std::min(g.nx, g.nz), std::min(g.nx, g.ny)));
  lw.q = static_cast<double*>(workspace(kWorkLblQ, g.cells() * sizeof(double)));
  lw.band = line_scratch(4, nmax, lw.L);
  lw.pst = static_cast<double*>(
      workspace(kWorkLblRes, kLblResWords * sizeof(double)));
  return (lw.q && lw.band && lw.pst) ? VK_OK : VK_ERR_HIP;
}

#define LBL_ARGS()                                                                \
  Geom g(*nx, *ny, *nz);                                                          \
  if (g.nx < 1 || g.ny < 1 || g.nz < 1)                                           \
    return fail(VK_ERR_BAD_SIZE, "linebyline: nx, ny, nz >= 1");                  \
  const size_t fb = g.cells() * sizeof(double);                                   \
  VK_RESOLVE(du, const double, u_temp, fb);                                       \
  VK_RESOLVE(dv, const double, v_temp, fb);                                       \
  VK_RESOLVE(dw, const double, w_temp, fb);                                       \
  VK_RESOLVE(dp, double, p, fb);                                                  \
  VK_RESOLVE(ddx, const double, dx, (g.nx + 2) * sizeof(double));                 \
  VK_RESOLVE(ddy, const double, dy, (g.ny + 2) * sizeof(double));                 \
  VK_RESOLVE(ddz, const double, dz, (g.nz + 2) * sizeof(double));                 \
  VK_RESOLVE(ddx12, const double, dx12, (g.nx + 2) * sizeof(double));             \
  VK_RESOLVE(ddy12, const double, dy12, (g.ny + 2) * sizeof(double));             \
  VK_RESOLVE(ddz12, const double, dz12, (g.nz + 2) * sizeof(double));             \
  const PMetrics pm{ddx, ddy, ddz, ddx12, ddy12, ddz12};                          \
  LblWork lw;                                                                     \
  if (lbl_work(g, lw)) return VK_ERR_HIP

int vk_linebyline(const int* nx, const int* ny, const int* nz, const double* u_temp,
                  const double* v_temp, const double* w_temp, const double* dx, const double* dy,
                  const double* dz, const double* dx12, const double* dy12, const double* dz12,
                  const double* dt, double* p) {
  VK_REQUIRE_INIT();
  LBL_ARGS();
  return linebyline(g, pm, du, dv, dw, *dt, dp, lw.q, lw.band, lw.L, lw.pst, rt().stream);
}

int vk_solve_p(double* p, const double* u_temp, const double* v_temp, const double* w_temp,
               const double* dt, const double* dx, const double* dy, const double* dz,
               const double* dx12, const double* dy12, const double* dz12, const int* nx,
               const int* ny, const int* nz, double* res, double* res_max, double* res_int,
               double* res_var, const int* max_iterations, int* iterations) {
  VK_REQUIRE_INIT();
  LBL_ARGS();
  if (!res || !res_max || !res_int || !res_var)
    return fail(VK_ERR_ARGUMENT, "solve_p: a residual output is null");
  const int nnn = max_iterations ? *max_iterations : 100000;
  if (nnn < 1) return fail(VK_ERR_ARGUMENT, "solve_p: max_iterations >= 1");
  hipStream_t s = rt().stream;
  double* part = lw.pst + kLblPart;
  double* out = lw.pst + kLblOut;
  const long long nint = (long long)g.interior();
  const int nb = (int)std::min<long long>(kResBlocks, (nint + 255) / 256);
  int done = 0;
  for (int lll = 1; lll <= nnn; lll++) {
    int rc = linebyline(g, pm, du, dv, dw, *dt, dp, lw.q, lw.band, lw.L, lw.pst, s);
    if (rc) return rc;
    k_lbl_resid<<<nb, 256, 0, s>>>(g, pm, du, dv, dw, dp, *dt, part);
    k_lbl_resid_final<<<1, 64, 0, s>>>(part, nb, out);
    double r[4];
    rc = hip_check(hipMemcpyAsync(r, out, sizeof(r), hipMemcpyDeviceToHost, s), "solve_p");
    if (!rc) rc = hip_check(hipStreamSynchronize(s), "solve_p");
    if (rc) return rc;
    *res = r[0];
    *res_max = r[1];
    *res_int = r[2];
    *res_var = r[3];
    done = lll;
    if (*res_max < 1.0e-005) break;
  }
  if (iterations) *iterations = done;
  return VK_OK;
}

}  // extern "C"
