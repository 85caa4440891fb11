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

// Multipliers recorded while factoring one reference line (a line whose
// matrix is the same for every line with no immersed solid: cellvalue 0,
// rhoCp = thc = 1).  With them the solve of such a line only applies the
// recorded multipliers to its right-hand side, performing exactly the
// floating-point operations of the general path on d.
struct TabLayout {
  int n, S, P, steps;
  long o_refa, o_refb, o_refc, o_f, o_g, o_h, o_afin, o_cfin, o_f0, o_k1, o_k2, o_fin0,
      o_fin1, o_fin2, total;
};
inline TabLayout tab_layout(int n, int S, bool cyclic) {
  TabLayout t;
  t.n = n;
  t.S = S;
  t.P = 2 * S;
  int steps = 0;
  for (int d = 1; d < (cyclic ? t.P / 2 : t.P); d <<= 1) steps++;
  t.steps = steps;
  long o = 0;
  t.o_refa = o; o += n;
  t.o_refb = o; o += n;
  t.o_refc = o; o += n;
  t.o_f = o; o += n;
  t.o_g = o; o += n;
  t.o_h = o; o += n;
  t.o_afin = o; o += n;
  t.o_cfin = o; o += n;
  t.o_f0 = o; o += S;
  t.o_k1 = o; o += (long)steps * t.P;
  t.o_k2 = o; o += (long)steps * t.P;
  t.o_fin0 = o; o += t.P;
  t.o_fin1 = o; o += t.P;
  t.o_fin2 = o; o += t.P;
  t.total = o;
  return t;
}

// One thread's segment: rows lo..lo+m-1 of its line (2 <= m <= MMAX).
// On return row r reads  a[r] x_first + x_r + c[r] x_last = d[r]  for the
// interior rows, and the first/last rows couple to the neighbouring
// segments' last/first unknowns through a[0] and c[m-1].  With REC the
// multipliers applied to d are stored into the table `tab`.
template <int MMAX, bool REC, typename V>
__device__ inline void segment_eliminate(int lo, int m, int seg, double (&a)[MMAX],
                                         double (&b)[MMAX], double (&c)[MMAX], V (&d)[MMAX],
                                         double* tab, const TabLayout& L) {
#pragma unroll
  for (int r = 0; r < MMAX; r++) {
    if (r < m) {
      if (r < 2) {
        double inv = 1.0 / b[r];
        if (REC) {
          tab[L.o_f + lo + r] = inv;
          tab[L.o_g + lo + r] = 0.0;
        }
        a[r] = a[r] * inv;
        c[r] = c[r] * inv;
        d[r] = inv * d[r];
      } else {
        double ar = a[r];
        double inv = 1.0 / (b[r] - ar * c[r - 1]);
        if (REC) {
          tab[L.o_f + lo + r] = inv;
          tab[L.o_g + lo + r] = ar;
        }
        d[r] = inv * (d[r] - ar * d[r - 1]);
        a[r] = -(inv * (ar * a[r - 1]));
        c[r] = inv * c[r];
      }
    }
  }
#pragma unroll
  for (int r = MMAX - 3; r >= 1; r--) {
    if (r <= m - 3) {
      if (REC) tab[L.o_h + lo + r] = c[r];
      d[r] = d[r] - c[r] * d[r + 1];
      a[r] = a[r] - c[r] * a[r + 1];
      c[r] = -(c[r] * c[r + 1]);
    }
  }
  if (m >= 3) {
    double inv = 1.0 / (1.0 - c[0] * a[1]);
    if (REC) {
      tab[L.o_f0 + seg] = inv;
      tab[L.o_h + lo] = c[0];
    }
    d[0] = inv * (d[0] - c[0] * d[1]);
    a[0] = inv * a[0];
    c[0] = -(inv * (c[0] * c[1]));
  }
  if (REC) {
#pragma unroll
    for (int r = 0; r < MMAX; r++)
      if (r < m) {
        tab[L.o_afin + lo + r] = a[r];
        tab[L.o_cfin + lo + r] = c[r];
      }
  }
}

// The same operations on d alone, with the recorded multipliers.
template <int MMAX, typename V>
__device__ inline void segment_apply(int lo, int m, int seg, V (&d)[MMAX], const double* tab,
                                     const TabLayout& L) {
#pragma unroll
  for (int r = 0; r < MMAX; r++) {
    if (r < m) {
      double f = tab[L.o_f + lo + r];
      if (r < 2)
        d[r] = f * d[r];
      else
        d[r] = f * (d[r] - tab[L.o_g + lo + r] * d[r - 1]);
    }
  }
#pragma unroll
  for (int r = MMAX - 3; r >= 1; r--)
    if (r <= m - 3) d[r] = d[r] - tab[L.o_h + lo + r] * d[r + 1];
  if (m >= 3) d[0] = tab[L.o_f0 + seg] * (d[0] - tab[L.o_h + lo] * d[1]);
}

// PCR on the P-row reduced system of one line held in LDS (row q has
// a[q] x_{q-1} + b[q] x_q + c[q] x_{q+1} = d[q]); this thread owns rows
// q0 and q0+1 and returns their solutions.  Every thread of the block
// runs the same number of barriers.
template <int P, bool REC, typename V>
__device__ inline void pcr_pair(double* la, double* lb, double* lc, LdsVec<V> ld, int q0,
                                bool cyclic, V& x0, V& x1, double* tab, const TabLayout& L) {
  const int nsteps_limit = cyclic ? P / 2 : P;
  int step = 0;
  for (int delta = 1; delta < nsteps_limit; delta <<= 1, step++) {
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
      if (REC) {
        tab[L.o_k1 + (long)step * P + q] = k1;
        tab[L.o_k2 + (long)step * P + q] = k2;
      }
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
      double inv = 1.0 / den;
      if (REC) {
        tab[L.o_fin0 + q] = inv;
        tab[L.o_fin1 + q] = bj;
        tab[L.o_fin2 + q] = ei;
      }
      xs[t] = inv * (bj * ld.get(q) - ei * ld.get(qq));
    } else {
      double inv = 1.0 / lb[q];
      if (REC) tab[L.o_fin0 + q] = inv;
      xs[t] = inv * ld.get(q);
    }
  }
  x0 = xs[0];
  x1 = xs[1];
}

// PCR on d alone with the recorded multipliers
template <int P, typename V>
__device__ inline void pcr_apply(LdsVec<V> ld, int q0, bool cyclic, V& x0, V& x1,
                                 const double* tab, const TabLayout& L) {
  const int nsteps_limit = cyclic ? P / 2 : P;
  int step = 0;
  for (int delta = 1; delta < nsteps_limit; delta <<= 1, step++) {
    V nd[2];
#pragma unroll
    for (int t = 0; t < 2; t++) {
      int q = q0 + t;
      int qm = q - delta, qp = q + delta;
      V dm, dp;
      vzero(dm);
      vzero(dp);
      if (cyclic) {
        qm &= (P - 1);
        qp &= (P - 1);
      }
      if (qm >= 0) dm = ld.get(qm);
      if (qp < P) dp = ld.get(qp);
      double k1 = tab[L.o_k1 + (long)step * P + q], k2 = tab[L.o_k2 + (long)step * P + q];
      nd[t] = ld.get(q) - k1 * dm - k2 * dp;
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < 2; t++) ld.set(q0 + t, nd[t]);
    __syncthreads();
  }
  V xs[2];
#pragma unroll
  for (int t = 0; t < 2; t++) {
    int q = q0 + t;
    if (cyclic) {
      int qq = (q + P / 2) & (P - 1);
      xs[t] = tab[L.o_fin0 + q] * (tab[L.o_fin1 + q] * ld.get(q) - tab[L.o_fin2 + q] * ld.get(qq));
    } else {
      xs[t] = tab[L.o_fin0 + q] * ld.get(q);
    }
  }
  x0 = xs[0];
  x1 = xs[1];
}

// Solve a batch of lines.  Sys provides
//   using V;  static constexpr bool kCyclic, kHasRef;
//   int n() const;                         rows per line
//   bool valid(int la, int lb) const;      line exists
//   void row(la, lb, r, a, b, c, d) const; row r of line (la, lb)
//   void row_ref(r, a, b, c) const;        row r of the reference line
//   void store(la, lb, r, V x) const;      solution of row r
// The line index is (la, lb): la = blockIdx.x * LPB + (line within block),
// lb = blockIdx.y.  When `tab` holds the reference line's multipliers and
// every row of every line of the block has the reference coefficients
// (bit for bit), the block applies the recorded multipliers instead of
// eliminating; otherwise it runs the general elimination.
template <int S, int MMAX, int NT, bool LANE_LINES, typename Sys>
__global__ __launch_bounds__(NT) void k_line_solve(Sys sys, const double* tab, TabLayout L) {
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

  double a[MMAX], b[MMAX], c[MMAX];
  V d[MMAX];
  bool clean = tab != nullptr;
#pragma unroll
  for (int r = 0; r < MMAX; r++) {
    a[r] = 0.0;
    b[r] = 1.0;
    c[r] = 0.0;
    vzero(d[r]);
    if (r < m && ok) {
      sys.row(la, lb, lo + r, a[r], b[r], c[r], d[r]);
      if (tab)
        clean = clean && a[r] == tab[L.o_refa + lo + r] && b[r] == tab[L.o_refb + lo + r] &&
                c[r] == tab[L.o_refc + lo + r];
    }
  }
  LdsVec<V> ld{s_d + line * P * vwords<V>()};
  const int q0 = 2 * seg;
  V xf, xl;
  const bool fast = __syncthreads_and(clean ? 1 : 0) != 0;
  if (fast) {
    segment_apply<MMAX, V>(lo, m, seg, d, tab, L);
    V dlast;
    vzero(dlast);
#pragma unroll
    for (int r = 0; r < MMAX; r++)
      if (r == m - 1) dlast = d[r];
    ld.set(q0, d[0]);
    ld.set(q0 + 1, dlast);
    __syncthreads();
    pcr_apply<P, V>(ld, q0, Sys::kCyclic, xf, xl, tab, L);
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
            x = d[r] - tab[L.o_afin + lo + r] * xf - tab[L.o_cfin + lo + r] * xl;
          sys.store(la, lb, lo + r, x);
        }
      }
    }
    return;
  }

  segment_eliminate<MMAX, false, V>(lo, m, seg, a, b, c, d, nullptr, L);
  double* la_ = s_a + line * P;
  double* lb_ = s_b + line * P;
  double* lc_ = s_c + line * P;
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
  pcr_pair<P, false, V>(la_, lb_, lc_, ld, q0, Sys::kCyclic, xf, xl, nullptr, L);
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

// Factor the reference line (one block, one thread per segment) and record
// its coefficients and multipliers into `tab`.
template <int S, int MMAX, typename Sys>
__global__ __launch_bounds__(S) void k_line_factor(Sys sys, double* tab, TabLayout L) {
  constexpr int P = 2 * S;
  __shared__ double s_a[P], s_b[P], s_c[P], s_d[P];
  const int seg = threadIdx.x;
  const int n = sys.n();
  const int lo = (seg * n) / S;
  const int m = ((seg + 1) * n) / S - lo;
  double a[MMAX], b[MMAX], c[MMAX], d[MMAX];
#pragma unroll
  for (int r = 0; r < MMAX; r++) {
    a[r] = 0.0;
    b[r] = 1.0;
    c[r] = 0.0;
    d[r] = 0.0;
    if (r < m) {
      sys.row_ref(lo + r, a[r], b[r], c[r]);
      tab[L.o_refa + lo + r] = a[r];
      tab[L.o_refb + lo + r] = b[r];
      tab[L.o_refc + lo + r] = c[r];
    }
  }
  segment_eliminate<MMAX, true, double>(lo, m, seg, a, b, c, d, tab, L);
  double alast = 0.0, clast = 0.0;
#pragma unroll
  for (int r = 0; r < MMAX; r++)
    if (r == m - 1) {
      alast = a[r];
      clast = c[r];
    }
  const int q0 = 2 * seg;
  s_a[q0] = a[0];
  s_b[q0] = 1.0;
  s_c[q0] = c[0];
  s_d[q0] = 0.0;
  s_a[q0 + 1] = alast;
  s_b[q0 + 1] = 1.0;
  s_c[q0 + 1] = clast;
  s_d[q0 + 1] = 0.0;
  __syncthreads();
  double x0, x1;
  pcr_pair<P, true, double>(s_a, s_b, s_c, LdsVec<double>{s_d}, q0, Sys::kCyclic, x0, x1, tab,
                            L);
}

// Host cache of reference-line tables, keyed by the system's description.
// A stale table is harmless: blocks whose rows do not match its recorded
// coefficients take the general path.
double* line_table(const void* key, size_t keylen, long total, bool* fresh);

// Launch helper: the segment count S is the smallest power of two (at
// most 64) that leaves at most 8 rows per segment (16 when 64 segments
// are not enough), so a thread keeps its rows in few registers and the
// reduced system has 2S rows, log2(2S) PCR steps.  x lines (segments
// across lanes) use 256-thread blocks; lines that are adjacent in memory
// (LANE_LINES) use at least 16 lines per block, so each row load of a
// wave covers whole 128-byte lines.
template <bool LANE_LINES, typename Sys>
inline int launch_line_solve(const Sys& sys, int n, int nla, int nlb, hipStream_t s) {
  int S = 1;
  while (S < 64 && (n + S - 1) / S > 8) S <<= 1;
  int mmax = (n + S - 1) / S;
  if (n < 2 || mmax > 16 || n < 2 * S) return -1;
  TabLayout L = tab_layout(n, S, Sys::kCyclic);
  double* tab = nullptr;
  bool fresh = false;
  if (Sys::kHasRef) {
    auto key = sys.ref_key(n, S);
    tab = line_table(&key, sizeof(key), L.total, &fresh);  // null: general path only
  }
#define VK_LS(SS, MM)                                                              \
  do {                                                                             \
    constexpr int NT_ = LANE_LINES ? (16 * SS > 512 ? 512 : (16 * SS < 256 ? 256 : 16 * SS)) \
                                   : 256;                                          \
    constexpr int LPB_ = NT_ / SS;                                                 \
    if (tab && fresh) k_line_factor<SS, MM, Sys><<<1, SS, 0, s>>>(sys, tab, L);    \
    dim3 grid((nla + LPB_ - 1) / LPB_, nlb);                                       \
    k_line_solve<SS, MM, NT_, LANE_LINES, Sys><<<grid, NT_, 0, s>>>(sys, tab, L);  \
    return 0;                                                                      \
  } while (0)
  if (mmax > 8) VK_LS(64, 16);
  switch (S) {
    case 64: VK_LS(64, 8);
    case 32: VK_LS(32, 8);
    case 16: VK_LS(16, 8);
    case 8: VK_LS(8, 8);
    case 4: VK_LS(4, 8);
    case 2: VK_LS(2, 8);
    default: VK_LS(1, 8);
  }
#undef VK_LS
}

}  // namespace vk
