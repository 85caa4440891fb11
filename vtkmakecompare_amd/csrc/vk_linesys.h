// vk_linesys.h -- the reference's implicit line systems as row functors
// for the partition solver of vk_partition.h, and the sweep launchers of
// solve_u_temp, solve_t and SPN.  The launchers are templates over the
// storage types; each storage combination is instantiated in a translation
// unit of its own (vk_sweep_*.hip) so that the library builds in parallel.
#pragma once

#include <string>
#include <type_traits>

#include "vk_kernels.h"
#include "vk_partition.h"
#include "vk_runtime.h"

namespace vk {

// ------------------------------------------------------ y line systems
//
// Row coefficients use the reciprocal metric arrays: the reference's
// -beta/alpha/dy12(j-1)/dy(j) becomes -(beta/alpha) * (1/dy12(j-1) *
// 1/dy(j)), and the diagonal 1 + beta/alpha*(...) is written 1 - a - c,
// the same quantity.  beta/alpha is exactly beta where cellvalue is 0.
// Each system computes a row from its per-point inputs with coef(); the
// reference row (cellvalue 0, rhoCp = thc = 1) is coef() on those inputs,
// which is what makes the recorded-multiplier path of vk_partition.h
// bit-identical to the general one.

// beta / alpha with alpha = 1 + dt eta cv; exactly beta where cv is 0 (the
// divisor is exactly 1), so no branch is needed
__device__ inline double beta_over_alpha(double beta, double dteta, double cv) {
  return beta / (dteta * cv + 1.0);
}

// key of a reference-line table: system kind, size, segmenting, the
// scalars and the caller's metric array the coefficients come from
struct RefKey {
  int kind, n, S, extra;
  double beta, dteta;
  const void* metric;
};

// harmonic face conductance (d1+d2)/(d1/k1+d2/k2) with one division
__device__ inline double hm(double d1, double k1, double d2, double k2) {
  return (d1 + d2) * k1 * k2 / (d1 * k2 + d2 * k1);
}

__device__ inline int clampi(int q, int lo, int hi) { return q < lo ? lo : (q > hi ? hi : q); }

// `v` where `keep`, else `other`, by bit masks: a select between a loaded
// value and a constant is turned into a branch around the load (which then
// waits for it before the join), so a thread's row loads would be issued
// one at a time
__device__ inline double keep_or(double v, bool keep, double other) {
  const unsigned long long m = 0ull - (unsigned long long)keep;
  return __longlong_as_double((__double_as_longlong(v) & (long long)m) |
                              (__double_as_longlong(other) & (long long)~m));
}

// per-row coefficient inputs: cellvalue alone, or cellvalue with the
// conductivity and heat capacity of the temperature systems
struct PtCv {
  double cv;
};
struct PtT {
  double cv, th, rc;
};
struct PtNone {};

// implicit_uy (nse_dominik.f:1505-1577) on line (i,k): rows j = 0..ny+1,
// ghost rows x0 + x1 = 0 and x_ny + x_{ny+1} = 0 (wall reflection); all
// rows are written back as parallel_uy does (nse_dominik.f:2907-2909).
template <typename TF = double, typename TC = double>
struct SysYFace {
  using V = double;
  using Pt = PtCv;
  static constexpr bool kCyclic = false, kHasRef = true, kNb = false;
  Geom g;
  TF* f;
  const TC* cv;
  RMetrics r;
  double beta, dteta;
  int ncv;
  const void* mkey;
  __device__ int n() const { return g.ny + 2; }
  __device__ bool valid(int la, int lb) const { return la < g.nx; }
  // ghost rows 0 and ny+1 are x0 + x1 = 0 and x_ny + x_{ny+1} = 0; the
  // interior expression is evaluated for every row (at a clamped j) so that
  // no metric load sits behind a branch
  __device__ void coef(int q, double cvv, double& a, double& b, double& c) const {
    const int j = clampi(q, 1, g.ny);
    const double rho = beta_over_alpha(beta, dteta, cvv);
    const double ai = -(rho * (r.rdy12[j - 1] * r.rdy[j]));
    const double ci = -(rho * (r.rdy12[j] * r.rdy[j]));
    const double bi = 1.0 - ai - ci;
    const bool in = q >= 1 && q <= g.ny;
    a = keep_or(ai, in, q == 0 ? 0.0 : 1.0);
    b = keep_or(bi, in, 1.0);
    c = keep_or(ci, in, q == 0 ? 1.0 : 0.0);
  }
  __device__ double rhs(int la, int lb, int q) const {
    double v = f[g.at32(la + 1, clampi(q, 1, g.ny), lb + 1)];
    return keep_or(v, q >= 1 && q <= g.ny, 0.0);
  }
  __device__ Pt pt(int la, int lb, int q) const {
    double v = cv[g.cv32(la + 1, clampi(q, 1, g.ny), lb + 1, ncv)];
    return Pt{keep_or(v, q >= 1 && q <= g.ny, 0.0)};
  }
  __device__ bool pt_ref(const Pt& p) const { return p.cv == 0.0; }
  __device__ void coef(int, int, int q, const Pt& p, const Pt&, const Pt&, double& a, double& b,
                       double& c) const {
    coef(q, p.cv, a, b, c);
  }
  __device__ void coef_ref(int q, double& a, double& b, double& c) const { coef(q, 0.0, a, b, c); }
  // one table per component: the components' sweeps run on streams of
  // their own, and each validates its table in its own stream order
  RefKey ref_key(int nn, int S) const { return RefKey{1, nn, S, ncv, beta, dteta, mkey}; }
  // element of row q, and whether the solve reads row q of its input
  __device__ unsigned addr(int la, int lb, int q) const { return g.at32(la + 1, q, lb + 1); }
  __device__ bool in_row(int q) const { return q >= 1 && q <= g.ny; }
  __device__ void store(int la, int lb, int q, double x) const {
    f[g.at32(la + 1, q, lb + 1)] = static_cast<TF>(x);
  }
  // dirty-line mask of this sweep (LineFlags): word la + lstride lb, bit
  // lbit; null when the sweep tests its points itself
  const unsigned* lflag = nullptr;
  unsigned lbit = 0;
  int lstride = 0;
};

// implicit_vy (nse_dominik.f:1583-1648): rows j = 0..ny with x0 = x_ny = 0
template <typename TF = double, typename TC = double>
struct SysYV {
  using V = double;
  using Pt = PtCv;
  static constexpr bool kCyclic = false, kHasRef = true, kNb = false;
  Geom g;
  TF* f;
  const TC* cv;
  RMetrics r;
  double beta, dteta;
  const void* mkey;
  __device__ int n() const { return g.ny + 1; }
  __device__ bool valid(int la, int lb) const { return la < g.nx; }
  __device__ void coef(int q, double cvv, double& a, double& b, double& c) const {
    const int j = clampi(q, 1, g.ny - 1);
    const double rho = beta_over_alpha(beta, dteta, cvv);
    const double ai = -(rho * (r.rdy[j] * r.rdy12[j]));
    const double ci = -(rho * (r.rdy[j + 1] * r.rdy12[j]));
    const double bi = 1.0 - ai - ci;
    const bool in = q >= 1 && q <= g.ny - 1;
    a = keep_or(ai, in, 0.0);
    b = keep_or(bi, in, 1.0);
    c = keep_or(ci, in, 0.0);
  }
  __device__ double rhs(int la, int lb, int q) const {
    double v = f[g.at32(la + 1, clampi(q, 1, g.ny - 1), lb + 1)];
    return keep_or(v, q >= 1 && q <= g.ny - 1, 0.0);
  }
  __device__ Pt pt(int la, int lb, int q) const {
    double v = cv[g.cv32(la + 1, clampi(q, 1, g.ny - 1), lb + 1, 3)];
    return Pt{keep_or(v, q >= 1 && q <= g.ny - 1, 0.0)};
  }
  __device__ bool pt_ref(const Pt& p) const { return p.cv == 0.0; }
  __device__ void coef(int, int, int q, const Pt& p, const Pt&, const Pt&, double& a, double& b,
                       double& c) const {
    coef(q, p.cv, a, b, c);
  }
  __device__ void coef_ref(int q, double& a, double& b, double& c) const { coef(q, 0.0, a, b, c); }
  RefKey ref_key(int nn, int S) const { return RefKey{2, nn, S, 3, beta, dteta, mkey}; }
  __device__ unsigned addr(int la, int lb, int q) const { return g.at32(la + 1, q, lb + 1); }
  __device__ bool in_row(int q) const { return q >= 1 && q <= g.ny - 1; }
  __device__ void store(int la, int lb, int q, double x) const {
    f[g.at32(la + 1, q, lb + 1)] = static_cast<TF>(x);
  }
  // dirty-line mask of this sweep (LineFlags): word la + lstride lb, bit
  // lbit; null when the sweep tests its points itself
  const unsigned* lflag = nullptr;
  unsigned lbit = 0;
  int lstride = 0;
};

// implicit_Ty (nse_dominik.f:1784-1870) with lly(0) = lly(1) and
// lly(ny+1) = lly(ny); only rows 1..ny are written back (parallel_Ty).
// The inputs of ghost rows 0 and ny+1 carry the wall cell's conductivity,
// which is the neighbour value rows 1 and ny use.
template <typename TF = double, typename TC = double>
struct SysYT {
  using V = double;
  using Pt = PtT;
  static constexpr bool kCyclic = false, kHasRef = true, kNb = true;
  Geom g;
  TF* f;
  const TC *cv, *rc, *thc;
  Metrics m;
  RMetrics r;
  double beta, dteta;
  TF* out = nullptr;  // where the solution goes (f itself when null)
  __device__ int n() const { return g.ny + 2; }
  __device__ bool valid(int la, int lb) const { return la < g.nx; }
  __device__ void coef(int q, double cvv, double l0, double lm, double lp, double rcv, double& a,
                       double& b, double& c) const {
    const int j = clampi(q, 1, g.ny);
    const double rho = beta_over_alpha(beta, dteta, cvv);
    const double rr = 1.0 / rcv;
    const double h1 = hm(m.dy[j], l0, m.dy[j - 1], lm);
    const double h2 = hm(m.dy[j + 1], lp, m.dy[j], l0);
    const double ai = -(rho * h1 * (r.rdy12[j - 1] * r.rdy[j]) * rr);
    const double ci = -(rho * h2 * (r.rdy12[j] * r.rdy[j]) * rr);
    const double bi = 1.0 - ai - ci;
    const bool in = q >= 1 && q <= g.ny;
    a = keep_or(ai, in, q == 0 ? 0.0 : 1.0);
    b = keep_or(bi, in, 1.0);
    c = keep_or(ci, in, q == 0 ? 1.0 : 0.0);
  }
  __device__ double rhs(int la, int lb, int q) const {
    double v = f[g.at32(la + 1, clampi(q, 1, g.ny), lb + 1)];
    return keep_or(v, q >= 1 && q <= g.ny, 0.0);
  }
  __device__ Pt pt(int la, int lb, int q) const {
    const int j = clampi(q, 1, g.ny), i = la + 1, k = lb + 1;
    const unsigned o = g.at32(i, j, k);
    const bool in = q >= 1 && q <= g.ny;
    double c = cv[g.cv32(i, j, k, 1)], t = thc[o], h = rc[o];
    return Pt{keep_or(c, in, 0.0), t, keep_or(h, in, 1.0)};
  }
  __device__ bool pt_ref(const Pt& p) const {
    return (p.cv == 0.0) & (p.th == 1.0) & (p.rc == 1.0);
  }
  __device__ void coef(int, int, int q, const Pt& p, const Pt& pm, const Pt& pp, double& a,
                       double& b, double& c) const {
    coef(q, p.cv, p.th, pm.th, pp.th, p.rc, a, b, c);
  }
  __device__ void coef_ref(int q, double& a, double& b, double& c) const {
    coef(q, 0.0, 1.0, 1.0, 1.0, 1.0, a, b, c);
  }
  RefKey ref_key(int nn, int S) const { return RefKey{3, nn, S, 1, beta, dteta, m.dy}; }
  __device__ unsigned addr(int la, int lb, int q) const { return g.at32(la + 1, q, lb + 1); }
  __device__ bool in_row(int q) const { return q >= 1 && q <= g.ny; }
  __device__ void store(int la, int lb, int q, double x) const {
    if (q >= 1 && q <= g.ny) (out ? out : f)[g.at32(la + 1, q, lb + 1)] = static_cast<TF>(x);
  }
  // dirty-line mask of this sweep (LineFlags): word la + lstride lb, bit
  // lbit; null when the sweep tests its points itself
  const unsigned* lflag = nullptr;
  unsigned lbit = 0;
  int lstride = 0;
};

// --------------------------------------------------- periodic systems

// A periodic line along axis `dir` (0 = x, 2 = z).  Lines are indexed
// (la, lb): for z lines (i, j) = (la+1, lb+1); for x lines (j, k) =
// (la+1, lb+1).  Row r is grid index q = r + 1.
struct LineAxis {
  Geom g;
  int dir;
  __device__ unsigned at(int la, int lb, int q) const {
    return dir == 2 ? g.at32(la + 1, lb + 1, q) : g.at32(q, la + 1, lb + 1);
  }
  __device__ unsigned cvat(int la, int lb, int q, int n) const {
    return dir == 2 ? g.cv32(la + 1, lb + 1, q, n) : g.cv32(q, la + 1, lb + 1, n);
  }
  __device__ int len() const { return dir == 2 ? g.nz : g.nx; }
};

// implicit_solver + cyclic (periodic_implicit.f:1-93) on rows q = 1..n;
// the ghost values the reference writes at q = 0 and n+1 (x0 = xn,
// x_{n+1} = x1) are written too.
template <typename TF = double, typename TC = double>
struct SysPeriodicFace {
  using V = double;
  using Pt = PtCv;
  static constexpr bool kCyclic = true, kHasRef = true, kNb = false;
  LineAxis ax;
  TF* f;
  const TC* cv;
  const double *rd, *rd12;
  double beta, dteta;
  int ncv, nla, nlb;
  const void* mkey;
  __device__ int n() const { return ax.len(); }
  __device__ bool valid(int la, int lb) const { return la < nla && lb < nlb; }
  __device__ void coef(int r, double cvv, double& a, double& b, double& c) const {
    int q = r + 1;
    double rho = beta_over_alpha(beta, dteta, cvv);
    a = -(rho * (rd[q] * rd12[q]));
    c = -(rho * (rd[q + 1] * rd12[q]));
    b = 1.0 - a - c;
  }
  __device__ double rhs(int la, int lb, int r) const { return f[ax.at(la, lb, r + 1)]; }
  __device__ Pt pt(int la, int lb, int r) const { return Pt{cv[ax.cvat(la, lb, r + 1, ncv)]}; }
  __device__ bool pt_ref(const Pt& p) const { return p.cv == 0.0; }
  __device__ void coef(int, int, int r, const Pt& p, const Pt&, const Pt&, double& a, double& b,
                       double& c) const {
    coef(r, p.cv, a, b, c);
  }
  __device__ void coef_ref(int r, double& a, double& b, double& c) const { coef(r, 0.0, a, b, c); }
  // one table per component (see SysYFace)
  RefKey ref_key(int nn, int S) const { return RefKey{4 + ax.dir, nn, S, ncv, beta, dteta, mkey}; }
  __device__ unsigned addr(int la, int lb, int r) const { return ax.at(la, lb, r + 1); }
  __device__ bool in_row(int) const { return true; }
  __device__ void store(int la, int lb, int r, double x) const {
    int q = r + 1, n = ax.len();
    const TF xs = static_cast<TF>(x);
    f[ax.at(la, lb, q)] = xs;
    if (q == n) f[ax.at(la, lb, 0)] = xs;
    if (q == 1) f[ax.at(la, lb, n + 1)] = xs;
  }
  // dirty-line mask of this sweep (LineFlags): word la + lstride lb, bit
  // lbit; null when the sweep tests its points itself
  const unsigned* lflag = nullptr;
  unsigned lbit = 0;
  int lstride = 0;
};

// implicit_Tx (nse_dominik.f:1725-1781), used along x and z by solve_t,
// with the conductivity ghosts wrapped periodically; rows 1..n written.
template <typename TF = double, typename TC = double>
struct SysPeriodicT {
  using V = double;
  using Pt = PtT;
  static constexpr bool kCyclic = true, kHasRef = true, kNb = true;
  LineAxis ax;
  TF* f;
  const TC *cv, *rc, *thc;
  const double *d, *rd, *rd12;
  double beta, dteta;
  int nla, nlb;
  __device__ int n() const { return ax.len(); }
  __device__ bool valid(int la, int lb) const { return la < nla && lb < nlb; }
  __device__ void coef(int r, double cvv, double l0, double lm, double lp, double rcv, double& a,
                       double& b, double& c) const {
    int q = r + 1;
    double rho = beta_over_alpha(beta, dteta, cvv);
    double rr = 1.0 / rcv;
    double h1 = hm(d[q], l0, d[q - 1], lm);
    double h2 = hm(d[q + 1], lp, d[q], l0);
    a = -(rho * h1 * (rd12[q - 1] * rd[q]) * rr);
    c = -(rho * h2 * (rd12[q] * rd[q]) * rr);
    b = 1.0 - a - c;
  }
  __device__ double rhs(int la, int lb, int r) const { return f[ax.at(la, lb, r + 1)]; }
  // r = -1 and r = n wrap to the last and first rows
  __device__ Pt pt(int la, int lb, int r) const {
    const int n = ax.len();
    const int q = r < 0 ? n : (r >= n ? 1 : r + 1);
    const unsigned o = ax.at(la, lb, q);
    return Pt{cv[ax.cvat(la, lb, q, 1)], thc[o], rc[o]};
  }
  __device__ bool pt_ref(const Pt& p) const {
    return (p.cv == 0.0) & (p.th == 1.0) & (p.rc == 1.0);
  }
  __device__ void coef(int, int, int r, const Pt& p, const Pt& pm, const Pt& pp, double& a,
                       double& b, double& c) const {
    coef(r, p.cv, p.th, pm.th, pp.th, p.rc, a, b, c);
  }
  __device__ void coef_ref(int r, double& a, double& b, double& c) const {
    coef(r, 0.0, 1.0, 1.0, 1.0, 1.0, a, b, c);
  }
  RefKey ref_key(int nn, int S) const { return RefKey{10 + ax.dir, nn, S, 1, beta, dteta, d}; }
  __device__ unsigned addr(int la, int lb, int r) const { return ax.at(la, lb, r + 1); }
  __device__ bool in_row(int) const { return true; }
  __device__ void store(int la, int lb, int r, double x) const {
    f[ax.at(la, lb, r + 1)] = static_cast<TF>(x);
  }
  // dirty-line mask of this sweep (LineFlags): word la + lstride lb, bit
  // lbit; null when the sweep tests its points itself
  const unsigned* lflag = nullptr;
  unsigned lbit = 0;
  int lstride = 0;
};

// -------------------------------------------------- spectral y solves

// SPN + CTRIDIAG (mkp.f:235-258, tridiag.f:43-58) for mode (I,K) =
// (la, lb), I = 0..nx/2, K = 0..nz-1, rows J = 1..ny with the Neumann
// fold of the first and last rows.  The zero mode is left to SP0B.  The
// matrix differs from mode to mode, so there is no reference line.
template <typename TS = double>
struct SysSpectral {
  using V = dcx;
  using Pt = PtNone;
  static constexpr bool kCyclic = false, kHasRef = false, kNb = false;
  static constexpr bool kWaves4 = true;  // k_line_general_w4
  static constexpr bool kBlock0 = true;  // SP0B's march, when block0 is set
  Geom g;
  TS* phit;  // interleaved complex
  const double *ac, *kxz;
  int block0 = 0;
  // SP0B (mkp.f:263-290) for the zero mode, which the other blocks leave
  // alone (store() skips it): the march of k_sp0b, in the solve's first
  // block.  The caller checks sp0b_fits_block0 for the scratch sizes.
  __device__ void run_block0(double* s1, int n1, double* s2, int n2) const {
    (void)n1;
    (void)n2;
    sp0b_march<TS>(g, ac, phit, s1, s2);
  }
  __device__ int n() const { return g.ny; }
  __device__ bool valid(int la, int lb) const { return la <= g.nx / 2 && lb < g.nz; }
  __device__ dcx rhs(int la, int lb, int r) const {
    unsigned o = 2 * g.spec32(la, r + 1, lb);
    return {phit[o], phit[o + 1]};
  }
  __device__ Pt pt(int, int, int) const { return Pt{}; }
  __device__ bool pt_ref(const Pt&) const { return false; }
  __device__ void coef(int la, int lb, int r, const Pt&, const Pt&, const Pt&, double& a,
                       double& b, double& c) const {
    const int J = r + 1, ny = g.ny;
    const double a1 = ac[J - 1], a3 = ac[J - 1 + 2 * ny];
    double a2 = ac[J - 1 + ny] - kxz[la + (g.nx / 2 + 1) * lb];
    // Neumann fold of the first and last rows, without a branch around the
    // loads
    const double f1 = a2 + a1, f3 = a2 + a3;
    a2 = keep_or(a2, J != 1, f1);
    a2 = keep_or(a2, J != ny, J == 1 ? f1 + a3 : f3);
    a = keep_or(a1, J != 1, 0.0);
    b = a2;
    c = keep_or(a3, J != ny, 0.0);
  }
  __device__ void coef_ref(int, double& a, double& b, double& c) const {
    a = 0.0;
    b = 1.0;
    c = 0.0;
  }
  RefKey ref_key(int nn, int S) const { return RefKey{0, nn, S, 0, 0.0, 0.0, nullptr}; }
  __device__ void store(int la, int lb, int r, dcx x) const {
    if (la == 0 && lb == 0) return;
    unsigned o = 2 * g.spec32(la, r + 1, lb);
    phit[o] = static_cast<TS>(x.x);
    phit[o + 1] = static_cast<TS>(x.y);
  }
  // dirty-line mask of this sweep (LineFlags): word la + lstride lb, bit
  // lbit; null when the sweep tests its points itself
  const unsigned* lflag = nullptr;
  unsigned lbit = 0;
  int lstride = 0;
};

// launch_line_solve's result as a status: -1 is a line length the solver
// does not take, -2 a failed allocation (whose message is already set)
static inline int line_status(int rc, const char* what) {
  if (rc == 0) return VK_OK;
  if (rc == -2) return VK_ERR_HIP;
  return fail(VK_ERR_BAD_SIZE, std::string(what) + ": line length not supported");
}

template <typename TS>
int launch_spn(const Geom& g, TS* phit, const double* ac, const double* kxz, hipStream_t s) {
  SysSpectral<TS> sys{g, phit, ac, kxz};
  ProfScope ps("line_spectral", s);
  return line_status(launch_line_solve<true>(sys, g.ny, g.nx / 2 + 1, g.nz, s), "spn");
}

// SPN with SP0B: the zero mode's march runs in the spectral solve's first
// block where its mode and coefficients fit the block's LDS arrays (512
// and 1024 doubles in every general-kernel shape of 256 threads), else
// as its own one-block kernel on an auxiliary stream beside the solve.
inline bool sp0b_fits_block0(const Geom& g) {
  return 2 * (g.ny + 2) <= 512 && 3 * g.ny <= 1024;
}
template <typename TS>
int launch_spn_sp0b(const Geom& g, TS* phit, const double* ac, const double* kxz, hipStream_t s) {
  if (!sp0b_fits_block0(g)) {
    AuxFork f0(s, 1u);
    if (f0.err) return f0.err;
    launch_sp0b<TS>(g, phit, ac, f0.aux(0));
    return launch_spn<TS>(g, phit, ac, kxz, s);
  }
  SysSpectral<TS> sys{g, phit, ac, kxz};
  sys.block0 = 1;
  ProfScope ps("line_spectral", s);
  return line_status(launch_line_solve<true>(sys, g.ny, g.nx / 2 + 1, g.nz, s), "spn");
}

// The transpose of a line system: row q of A^T holds (c_{q-1}, b_q,
// a_{q+1}) of A (indices mod n when cyclic, zero past the ends
// otherwise).  The right-hand side is the adjoint of the solve's output,
// read from `lam` on the rows the forward solve reads from its input (the
// others are constant rows of the forward system, zero here); the solution
// overwrites `lam` on those rows.
//
// A line whose forward rows all equal the reference line's has the
// transposed reference rows, so the recorded-multiplier path applies to
// the transpose too, with its own table (coef_ref below).  Which lines are
// clean comes from the forward sweep's dirty-line mask (lflag); without
// one every line is taken as dirty (pt_ref is false) and solved by the
// general kernel.
template <typename S>
struct SysT {
  using V = double;
  using Pt = PtNone;
  static constexpr bool kCyclic = S::kCyclic, kHasRef = S::kHasRef, kNb = false;
  S s;
  double* lam;
  __device__ int n() const { return s.n(); }
  __device__ bool valid(int la, int lb) const { return s.valid(la, lb); }
  __device__ void row(int la, int lb, int q, double& a, double& b, double& c) const {
    typename S::Pt p = s.pt(la, lb, q);
    typename S::Pt pm = S::kNb ? s.pt(la, lb, q - 1) : typename S::Pt();
    typename S::Pt pp = S::kNb ? s.pt(la, lb, q + 1) : typename S::Pt();
    s.coef(la, lb, q, p, pm, pp, a, b, c);
  }
  __device__ double rhs(int la, int lb, int q) const {
    const double v = lam[s.addr(la, lb, q)];
    return s.in_row(q) ? v : 0.0;
  }
  __device__ Pt pt(int, int, int) const { return Pt{}; }
  __device__ bool pt_ref(const Pt&) const { return false; }
  __device__ void coef(int la, int lb, int q, const Pt&, const Pt&, const Pt&, double& a,
                       double& b, double& c) const {
    const int nn = s.n();
    double a0, b0, c0, am, bm, cm, ap, bp, cp;
    row(la, lb, q, a0, b0, c0);
    if (kCyclic) {
      row(la, lb, q == 0 ? nn - 1 : q - 1, am, bm, cm);
      row(la, lb, q == nn - 1 ? 0 : q + 1, ap, bp, cp);
      a = cm;
      c = ap;
    } else {
      a = 0.0;
      c = 0.0;
      if (q >= 1) {
        row(la, lb, q - 1, am, bm, cm);
        a = cm;
      }
      if (q <= nn - 2) {
        row(la, lb, q + 1, ap, bp, cp);
        c = ap;
      }
    }
    b = b0;
  }
  // row q of the transposed reference line, as coef transposes row()
  __device__ void coef_ref(int q, double& a, double& b, double& c) const {
    const int nn = s.n();
    double a0, b0, c0, am, bm, cm, ap, bp, cp;
    s.coef_ref(q, a0, b0, c0);
    if (kCyclic) {
      s.coef_ref(q == 0 ? nn - 1 : q - 1, am, bm, cm);
      s.coef_ref(q == nn - 1 ? 0 : q + 1, ap, bp, cp);
      a = cm;
      c = ap;
    } else {
      a = 0.0;
      c = 0.0;
      if (q >= 1) {
        s.coef_ref(q - 1, am, bm, cm);
        a = cm;
      }
      if (q <= nn - 2) {
        s.coef_ref(q + 1, ap, bp, cp);
        c = ap;
      }
    }
    b = b0;
  }
  // the forward system's key, marked as its transpose
  RefKey ref_key(int nn, int S_) const {
    RefKey k = s.ref_key(nn, S_);
    k.kind += 100;
    return k;
  }
  __device__ void store(int la, int lb, int q, double x) const {
    if (s.in_row(q)) lam[s.addr(la, lb, q)] = x;
  }
  // dirty-line mask of this sweep (LineFlags): word la + lstride lb, bit
  // lbit; null when the sweep tests its points itself
  const unsigned* lflag = nullptr;
  unsigned lbit = 0;
  int lstride = 0;
};

// Runs fn(sys, n, nla, nlb, lane_lines, name) on the system of one sweep
// of solve_u_temp: comp 0,1,2 = u, v, w (cellvalue slot comp+2), dir 0, 1,
// 2 = the y, z, x sweep; v's z and x sweeps stop at j = ny-1
// (nse_dominik.f:1411,1418).
template <typename Sys>
inline void use_flags(Sys& sys, const unsigned* words, unsigned bit, int stride) {
  sys.lflag = words;
  sys.lbit = bit;
  sys.lstride = stride;
}

template <typename TC, typename Fn>
int with_velocity_system(const Geom& g, const Metrics& m, const RMetrics& r, double* f,
                         const TC* cv, int comp, int dir, double beta, double dteta, Fn&& fn,
                         const LineFlags* fl = nullptr) {
  const int jmax = (comp == 1) ? g.ny - 1 : g.ny;
  const unsigned bit = 1u << (comp + 1);
  if (dir == 0) {
    if (comp == 1) {
      SysYV<double, TC> sy{g, f, cv, r, beta, dteta, m.dy};
      if (fl) use_flags(sy, fl->fy, bit, g.nx);
      return fn(sy, g.ny + 1, g.nx, g.nz, std::true_type(), "line_y_v");
    }
    SysYFace<double, TC> sy{g, f, cv, r, beta, dteta, comp + 2, m.dy};
    if (fl) use_flags(sy, fl->fy, bit, g.nx);
    return fn(sy, g.ny + 2, g.nx, g.nz, std::true_type(), "line_y_face");
  }
  if (dir == 1) {
    SysPeriodicFace<double, TC> sz{LineAxis{g, 2}, f,    cv,   r.rdz, r.rdz12, beta,
                                   dteta,          comp + 2, g.nx, jmax,  m.dz};
    if (fl) use_flags(sz, fl->fz, bit, g.nx);
    return fn(sz, g.nz, g.nx, jmax, std::true_type(), "line_z_face");
  }
  SysPeriodicFace<double, TC> sx{LineAxis{g, 0}, f,    cv,   r.rdx, r.rdx12, beta,
                                 dteta,          comp + 2, jmax, g.nz,  m.dx};
  if (fl) use_flags(sx, fl->fx, bit, g.ny);
  return fn(sx, g.nx, jmax, g.nz, std::false_type(), "line_x_face");
}

// the same for solve_t: dir 0, 1, 2 = the y, x, z sweep (nse_dominik.f:1477-1500)
template <typename TT, typename Fn>
int with_temperature_system(const Geom& g, const Metrics& m, const RMetrics& r, TT* f,
                            const TT* cv, const TT* rc_, const TT* thc, int dir, double beta,
                            double dteta, Fn&& fn, TT* out = nullptr,
                            const LineFlags* fl = nullptr) {
  if (dir == 0) {
    SysYT<TT, TT> sy{g, f, cv, rc_, thc, m, r, beta, dteta, out};
    if (fl) use_flags(sy, fl->fy, 1u, g.nx);
    return fn(sy, g.ny + 2, g.nx, g.nz, std::true_type(), "line_y_t");
  }
  if (dir == 1) {
    SysPeriodicT<TT, TT> sx{LineAxis{g, 0}, f, cv, rc_, thc, m.dx, r.rdx, r.rdx12, beta, dteta,
                            g.ny, g.nz};
    if (fl) use_flags(sx, fl->fx, 1u, g.ny);
    return fn(sx, g.nx, g.ny, g.nz, std::false_type(), "line_x_t");
  }
  SysPeriodicT<TT, TT> sz{LineAxis{g, 2}, f, cv, rc_, thc, m.dz, r.rdz, r.rdz12, beta, dteta,
                          g.nx, g.ny};
  if (fl) use_flags(sz, fl->fz, 1u, g.nx);
  return fn(sz, g.nz, g.nx, g.ny, std::true_type(), "line_z_t");
}

// a forward solve of the system handed over by with_*_system
struct SolveFwd {
  hipStream_t s;
  template <typename Sys, typename Lane>
  int operator()(const Sys& sys, int n, int nla, int nlb, Lane, const char* name) const {
    ProfScope ps(name, s);
    return line_status(launch_line_solve<Lane::value>(sys, n, nla, nlb, s), name);
  }
};

// the transposed solve on the adjoint field `lam`
struct SolveAdj {
  hipStream_t s;
  double* lam;
  template <typename Sys, typename Lane>
  int operator()(const Sys& sys, int n, int nla, int nlb, Lane, const char* name) const {
    SysT<Sys> st{sys, lam};
    // the forward sweep's dirty-line mask, when the system carries one
    use_flags(st, sys.lflag, sys.lbit, sys.lstride);
    ProfScope ps("line_adjoint", s);
    return line_status(launch_line_solve<Lane::value>(st, n, nla, nlb, s), name);
  }
};

template <typename TC>
int launch_velocity_sweep(const Geom& g, const Metrics& m, const RMetrics& r, double* f,
                          const TC* cv, int comp, int dir, double beta, double dteta,
                          hipStream_t s, const LineFlags* fl) {
  return with_velocity_system(g, m, r, f, cv, comp, dir, beta, dteta, SolveFwd{s}, fl);
}

template <typename TT>
int launch_temperature_sweep(const Geom& g, const Metrics& m, const RMetrics& r, TT* f,
                             const TT* cv, const TT* rc_, const TT* thc, int dir, double beta,
                             double dteta, hipStream_t s, TT* out, const LineFlags* fl) {
  return with_temperature_system(g, m, r, f, cv, rc_, thc, dir, beta, dteta, SolveFwd{s}, out,
                                 fl);
}

// solve_u_temp body after the penalisation divide: y, z, x sweeps of one
// component
template <typename TC>
int launch_adi_velocity(const Geom& g, const Metrics& m, const RMetrics& r, double* f,
                        const TC* cv, int comp, double beta, double dteta, hipStream_t s,
                        const LineFlags* fl) {
  for (int dir = 0; dir < 3; dir++) {
    int rc = launch_velocity_sweep<TC>(g, m, r, f, cv, comp, dir, beta, dteta, s, fl);
    if (rc) return rc;
  }
  return VK_OK;
}

// solve_t: y, x, z sweeps; with `out` the y sweep writes the interior of
// `out` and the x and z sweeps work there (f is then only read)
template <typename TT>
int launch_adi_temperature(const Geom& g, const Metrics& m, const RMetrics& r, TT* f,
                           const TT* cv, const TT* rc_, const TT* thc, double beta, double dteta,
                           hipStream_t s, TT* out, const LineFlags* fl) {
  for (int dir = 0; dir < 3; dir++) {
    TT* src = (dir == 0 || !out) ? f : out;
    int rc = launch_temperature_sweep<TT>(g, m, r, src, cv, rc_, thc, dir, beta, dteta, s,
                                          dir == 0 ? out : nullptr, fl);
    if (rc) return rc;
  }
  return VK_OK;
}

}  // namespace vk
