// vk_partition.h -- batched tridiagonal line solver for gfx950.
//
// Every implicit line solve of the hot path (the ADI sweeps of
// solve_u_temp/solve_t and the spectral y solves of SPN) is a batch of
// independent tridiagonal systems.  Each line of n rows is split into S
// segments (S a power of two, one thread each).  A thread loads its rows
// straight from HBM, eliminates inside the segment in registers so that
// every row couples only to the segment's first and last unknowns (the
// partition method of Laszlo, Giles and Appleyard, ACM TOMS 42(4), 2016),
// and writes the segment's two boundary rows to LDS.  The 2S-row reduced
// system is solved by parallel cyclic reduction in LDS; a periodic line is
// a cyclic reduced system, which PCR closes with a 2x2 solve, so no
// Sherman-Morrison correction and no second right-hand side is needed.
// Each thread then back-substitutes its rows and stores them.  HBM traffic
// is one read of the coefficients' inputs and the right-hand side and one
// write of the solution; nothing is staged in global scratch.
//
// Two thread layouts:
//   LANE_LINES: consecutive lanes take consecutive lines (lines that are
//     adjacent in memory, e.g. y or z lines indexed by i), so a row load
//     by a wave reads contiguous 8-double groups per segment;
//   otherwise consecutive lanes take consecutive segments of one line
//     (x lines, which are contiguous in memory).
#pragma once

#include "vk_common.h"

namespace vk {

struct dcx {
  double x, y;
};
__device__ inline dcx operator+(dcx a, dcx b) { return {a.x + b.x, a.y + b.y}; }
__device__ inline dcx operator-(dcx a, dcx b) { return {a.x - b.x, a.y - b.y}; }
__device__ inline dcx operator*(double s, dcx v) { return {s * v.x, s * v.y}; }
__device__ inline dcx operator-(dcx a) { return {-a.x, -a.y}; }
__device__ inline void vzero(double& v) { v = 0.0; }
__device__ inline void vzero(dcx& v) { v = {0.0, 0.0}; }

template <typename V>
struct LdsVec;
template <>
struct LdsVec<double> {
  double* p;
  __device__ double get(int q) const { return p[q]; }
  __device__ void set(int q, double v) const { p[q] = v; }
};
template <>
struct LdsVec<dcx> {
  double* p;  // re at [2q], im at [2q+1]
  __device__ dcx get(int q) const { return {p[2 * q], p[2 * q + 1]}; }
  __device__ void set(int q, dcx v) const {
    p[2 * q] = v.x;
    p[2 * q + 1] = v.y;
  }
};
template <typename V>
constexpr int vwords() {
  return sizeof(V) / sizeof(double);
}

// One thread's segment: rows lo..lo+m-1 of its line (2 <= m <= MMAX).
// On return row r reads  a[r] x_first + x_r + c[r] x_last = d[r]  for the
// interior rows, and the first/last rows couple to the neighbouring
// segments' last/first unknowns through a[0] and c[m-1].
template <int MMAX, typename V, typename Sys>
__device__ inline void segment_eliminate(const Sys& sys, int la, int lb, bool ok, int lo, int m,
                                         double (&a)[MMAX], double (&c)[MMAX], V (&d)[MMAX]) {
#pragma unroll
  for (int r = 0; r < MMAX; r++) {
    if (r < m) {
      double ar = 0.0, br = 1.0, cr = 0.0;
      V dr;
      vzero(dr);
      if (ok) sys.row(la, lb, lo + r, ar, br, cr, dr);
      if (r < 2) {
        double inv = 1.0 / br;
        a[r] = ar * inv;
        c[r] = cr * inv;
        d[r] = inv * dr;
      } else {
        double inv = 1.0 / (br - ar * c[r - 1]);
        d[r] = inv * (dr - ar * d[r - 1]);
        a[r] = -(inv * (ar * a[r - 1]));
        c[r] = inv * cr;
      }
    }
  }
#pragma unroll
  for (int r = MMAX - 3; r >= 1; r--) {
    if (r <= m - 3) {
      d[r] = d[r] - c[r] * d[r + 1];
      a[r] = a[r] - c[r] * a[r + 1];
      c[r] = -(c[r] * c[r + 1]);
    }
  }
  if (m >= 3) {
    double inv = 1.0 / (1.0 - c[0] * a[1]);
    d[0] = inv * (d[0] - c[0] * d[1]);
    a[0] = inv * a[0];
    c[0] = -(inv * (c[0] * c[1]));
  }
}

// PCR on the P-row reduced system of one line held in LDS (row q has
// a[q] x_{q-1} + b[q] x_q + c[q] x_{q+1} = d[q]); this thread owns rows
// q0 and q0+1 and returns their solutions.  Every thread of the block
// runs the same number of barriers.
template <int P, typename V>
__device__ inline void pcr_pair(double* la, double* lb, double* lc, LdsVec<V> ld, int q0,
                                bool cyclic, V& x0, V& x1) {
  const int nsteps_limit = cyclic ? P / 2 : P;
  for (int delta = 1; delta < nsteps_limit; delta <<= 1) {
    double na[2], nb[2], nc[2];
    V nd[2];
#pragma unroll
    for (int t = 0; t < 2; t++) {
      int q = q0 + t;
      int qm = q - delta, qp = q + delta;
      double am = 0.0, bm = 1.0, cm = 0.0, ap = 0.0, bp = 1.0, cp = 0.0;
      V dm, dp;
      vzero(dm);
      vzero(dp);
      if (cyclic) {
        qm &= (P - 1);
        qp &= (P - 1);
      }
      if (qm >= 0) {
        am = la[qm]; bm = lb[qm]; cm = lc[qm]; dm = ld.get(qm);
      }
      if (qp < P) {
        ap = la[qp]; bp = lb[qp]; cp = lc[qp]; dp = ld.get(qp);
      }
      double k1 = la[q] / bm, k2 = lc[q] / bp;
      na[t] = -(am * k1);
      nc[t] = -(cp * k2);
      nb[t] = lb[q] - cm * k1 - ap * k2;
      nd[t] = ld.get(q) - k1 * dm - k2 * dp;
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < 2; t++) {
      la[q0 + t] = na[t];
      lb[q0 + t] = nb[t];
      lc[q0 + t] = nc[t];
      ld.set(q0 + t, nd[t]);
    }
    __syncthreads();
  }
  V xs[2];
#pragma unroll
  for (int t = 0; t < 2; t++) {
    int q = q0 + t;
    if (cyclic) {
      int qq = (q + P / 2) & (P - 1);
      double ei = la[q] + lc[q], ej = la[qq] + lc[qq];
      double bi = lb[q], bj = lb[qq];
      double den = bi * bj - ei * ej;
      xs[t] = (1.0 / den) * (bj * ld.get(q) - ei * ld.get(qq));
    } else {
      xs[t] = (1.0 / lb[q]) * ld.get(q);
    }
  }
  x0 = xs[0];
  x1 = xs[1];
}

// Solve a batch of lines.  Sys provides
//   using V;  static constexpr bool kCyclic;
//   int n() const;                         rows per line
//   bool valid(int la, int lb) const;      line exists
//   void row(la, lb, r, a, b, c, d) const; row r of line (la, lb)
//   void store(la, lb, r, V x) const;      solution of row r
// The line index is (la, lb): la = blockIdx.x * LPB + (line within block),
// lb = blockIdx.y.
template <int S, int MMAX, int NT, bool LANE_LINES, typename Sys>
__global__ __launch_bounds__(NT) void k_line_solve(Sys sys) {
  using V = typename Sys::V;
  constexpr int LPB = NT / S;
  constexpr int P = 2 * S;
  __shared__ double s_a[LPB * P], s_b[LPB * P], s_c[LPB * P];
  __shared__ double s_d[LPB * P * vwords<V>()];
  const int tid = threadIdx.x;
  const int line = LANE_LINES ? (tid % LPB) : (tid / S);
  const int seg = LANE_LINES ? (tid / LPB) : (tid % S);
  const int la = blockIdx.x * LPB + line, lb = blockIdx.y;
  const bool ok = sys.valid(la, lb);
  const int n = sys.n();
  const int lo = (seg * n) / S;
  const int m = ((seg + 1) * n) / S - lo;

  double a[MMAX], c[MMAX];
  V d[MMAX];
  segment_eliminate<MMAX, V>(sys, la, lb, ok, lo, m, a, c, d);

  double* la_ = s_a + line * P;
  double* lb_ = s_b + line * P;
  double* lc_ = s_c + line * P;
  LdsVec<V> ld{s_d + line * P * vwords<V>()};
  const int q0 = 2 * seg;
  double alast = 0.0, clast = 0.0;
  V dlast;
  vzero(dlast);
#pragma unroll
  for (int r = 0; r < MMAX; r++)
    if (r == m - 1) {
      alast = a[r];
      clast = c[r];
      dlast = d[r];
    }
  la_[q0] = a[0];
  lb_[q0] = 1.0;
  lc_[q0] = c[0];
  ld.set(q0, d[0]);
  la_[q0 + 1] = alast;
  lb_[q0 + 1] = 1.0;
  lc_[q0 + 1] = clast;
  ld.set(q0 + 1, dlast);
  __syncthreads();

  V xf, xl;
  pcr_pair<P, V>(la_, lb_, lc_, ld, q0, Sys::kCyclic, xf, xl);

  if (ok) {
#pragma unroll
    for (int r = 0; r < MMAX; r++) {
      if (r < m) {
        V x;
        if (r == 0)
          x = xf;
        else if (r == m - 1)
          x = xl;
        else
          x = d[r] - a[r] * xf - c[r] * xl;
        sys.store(la, lb, lo + r, x);
      }
    }
  }
}

// Launch helper: picks the segment count (a power of two with at least two
// rows per segment) and the register bound of rows per segment.
template <bool LANE_LINES, typename Sys>
inline int launch_line_solve(const Sys& sys, int n, int nla, int nlb, hipStream_t s) {
  int S = 1;
  while (S < 64 && 4 * S <= n) S <<= 1;  // S <= n/2 and S <= 64
  int mmax = (n + S - 1) / S;
  if (n < 2 || mmax > 16) return -1;
#define VK_LS(SS, MM, NTT)                                                         \
  do {                                                                             \
    constexpr int LPB_ = NTT / SS;                                                 \
    dim3 grid((nla + LPB_ - 1) / LPB_, nlb);                                       \
    k_line_solve<SS, MM, NTT, LANE_LINES, Sys><<<grid, NTT, 0, s>>>(sys);          \
    return 0;                                                                      \
  } while (0)
  constexpr int NTL = LANE_LINES ? 512 : 256;
  if (S == 64) {
    if (mmax <= 4) VK_LS(64, 4, NTL);
    if (mmax <= 8) VK_LS(64, 8, NTL);
    VK_LS(64, 16, NTL);
  }
  if (mmax > 4) return -1;
  switch (S) {
    case 32: VK_LS(32, 4, NTL);
    case 16: VK_LS(16, 4, NTL);
    case 8: VK_LS(8, 4, NTL);
    case 4: VK_LS(4, 4, NTL);
    case 2: VK_LS(2, 4, NTL);
    default: VK_LS(1, 4, NTL);
  }
#undef VK_LS
}

}  // namespace vk
