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

#include <type_traits>

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

// A vector of reduced right-hand sides in LDS: real values at p[i]; complex
// ones with the imaginary parts one plane (`plane` doubles) further on, so
// that a lane's access is one 8-byte word.
template <typename V>
struct LdsVec;
template <>
struct LdsVec<double> {
  double* p;
  int plane;
  __device__ double get(int i) const { return p[i]; }
  __device__ void set(int i, double v) const { p[i] = v; }
};
template <>
struct LdsVec<dcx> {
  double* p;
  int plane;
  __device__ dcx get(int i) const { return {p[i], p[i + plane]}; }
  __device__ void set(int i, dcx v) const {
    p[i] = v.x;
    p[i + plane] = v.y;
  }
};

// Where row q of a line's 2S-row reduced system sits in LDS: slot
// (q & 1) * S + (q >> 1) -- the first rows of all segments, then their last
// rows -- times `stride` past `base`.  With lines interleaved (stride = the
// block's line count) the lanes of a wave reach consecutive words for their
// own rows and for every PCR neighbour, free of bank conflicts.
template <int S>
struct LSlot {
  int base, stride;
  __device__ int at(int q) const { return base + ((q & 1) * S + (q >> 1)) * stride; }
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

// PCR on the P = 2S-row reduced system of one line held in LDS (row q has
// a[q] x_{q-1} + b[q] x_q + c[q] x_{q+1} = d[q]); the thread of segment `seg`
// owns rows 2 seg and 2 seg + 1 and returns their solutions.  Every thread
// of the block runs the same number of barriers.  With REC the multipliers
// applied to d are recorded in `tab` by natural row index.
template <int S, bool REC, typename V>
__device__ inline void pcr_pair(double* la, double* lb, double* lc, LdsVec<V> ld, LSlot<S> sl,
                                int seg, bool cyclic, V& x0, V& x1, double* tab,
                                const TabLayout& L) {
  constexpr int P = 2 * S;
  const int q0 = 2 * seg;
  const int nsteps_limit = cyclic ? P / 2 : P;
  int step = 0;
  for (int delta = 1; delta < nsteps_limit; delta <<= 1, step++) {
    double na[2], nb[2], nc[2];
    V nd[2];
#pragma unroll
    for (int t = 0; t < 2; t++) {
      const int q = q0 + t;
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
        const int i = sl.at(qm);
        am = la[i]; bm = lb[i]; cm = lc[i]; dm = ld.get(i);
      }
      if (qp < P) {
        const int i = sl.at(qp);
        ap = la[i]; bp = lb[i]; cp = lc[i]; dp = ld.get(i);
      }
      const int iq = sl.at(q);
      double k1 = la[iq] / bm, k2 = lc[iq] / bp;
      if (REC) {
        tab[L.o_k1 + (long)step * P + q] = k1;
        tab[L.o_k2 + (long)step * P + q] = k2;
      }
      na[t] = -(am * k1);
      nc[t] = -(cp * k2);
      nb[t] = lb[iq] - cm * k1 - ap * k2;
      nd[t] = ld.get(iq) - k1 * dm - k2 * dp;
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < 2; t++) {
      const int iq = sl.at(q0 + t);
      la[iq] = na[t];
      lb[iq] = nb[t];
      lc[iq] = nc[t];
      ld.set(iq, nd[t]);
    }
    __syncthreads();
  }
  V xs[2];
#pragma unroll
  for (int t = 0; t < 2; t++) {
    const int q = q0 + t;
    const int iq = sl.at(q);
    if (cyclic) {
      const int iqq = sl.at((q + P / 2) & (P - 1));
      double ei = la[iq] + lc[iq], ej = la[iqq] + lc[iqq];
      double bi = lb[iq], bj = lb[iqq];
      double den = bi * bj - ei * ej;
      double inv = 1.0 / den;
      if (REC) {
        tab[L.o_fin0 + q] = inv;
        tab[L.o_fin1 + q] = bj;
        tab[L.o_fin2 + q] = ei;
      }
      xs[t] = inv * (bj * ld.get(iq) - ei * ld.get(iqq));
    } else {
      double inv = 1.0 / lb[iq];
      if (REC) tab[L.o_fin0 + q] = inv;
      xs[t] = inv * ld.get(iq);
    }
  }
  x0 = xs[0];
  x1 = xs[1];
}

// The reference line's table as staged in LDS by the fast kernel.  The
// per-row parts (f, g, h, afin, cfin) are laid out by (row within segment,
// segment), the per-reduced-row parts (k1, k2 per PCR step, fin0..2) by
// the reduced-row slot (q & 1) * S + (q >> 1), so the lanes of a wave read
// consecutive or equal words.
template <int S, int MMAX>
struct LdsTab {
  static constexpr int kRow = MMAX * S;
  static constexpr int P = 2 * S;
  int steps, f, g, h, afin, cfin, f0, k1, k2, fin0, fin1, fin2, total;
  __host__ __device__ explicit LdsTab(int nsteps) : steps(nsteps) {
    f = 0;
    g = f + kRow;
    h = g + kRow;
    afin = h + kRow;
    cfin = afin + kRow;
    f0 = cfin + kRow;
    k1 = f0 + S;
    k2 = k1 + steps * P;
    fin0 = k2 + steps * P;
    fin1 = fin0 + P;
    fin2 = fin1 + P;
    total = fin2 + P;
  }
  // global table index of LDS slot u (n rows per line); slots of rows past
  // a short segment map to the segment's first row
  __device__ long src(int u, int n, const TabLayout& L) const {
    if (u < f0) {
      const int part = u / kRow, rem = u % kRow;
      const int r = rem / S, seg = rem % S;
      const int lo = (seg * n) / S, m = ((seg + 1) * n) / S - lo;
      const long base = part == 0 ? L.o_f : part == 1 ? L.o_g : part == 2 ? L.o_h
                        : part == 3 ? L.o_afin : L.o_cfin;
      return base + lo + (r < m ? r : 0);
    }
    if (u < k1) return L.o_f0 + (u - f0);
    int v, base;
    if (u < k2) {
      v = u - k1;
      base = (int)L.o_k1;
    } else if (u < fin0) {
      v = u - k2;
      base = (int)L.o_k2;
    } else if (u < fin1) {
      v = u - fin0;
      base = (int)L.o_fin0;
    } else if (u < fin2) {
      v = u - fin1;
      base = (int)L.o_fin1;
    } else {
      v = u - fin2;
      base = (int)L.o_fin2;
    }
    const int step = v / P, w = v % P;
    const int t = w / S, seg = w % S;
    return (long)base + (long)step * P + 2 * seg + t;
  }
};

// PCR on d alone with the recorded multipliers of the LDS table
template <int S, int MMAX, typename V>
__device__ inline void pcr_apply(LdsVec<V> ld, LSlot<S> sl, int seg, bool cyclic, V& x0, V& x1,
                                 const double* st, const LdsTab<S, MMAX>& T) {
  constexpr int P = 2 * S;
  const int q0 = 2 * seg;
  const int nsteps_limit = cyclic ? P / 2 : P;
  int step = 0;
  for (int delta = 1; delta < nsteps_limit; delta <<= 1, step++) {
    V nd[2];
#pragma unroll
    for (int t = 0; t < 2; t++) {
      const int q = q0 + t;
      int qm = q - delta, qp = q + delta;
      V dm, dp;
      vzero(dm);
      vzero(dp);
      if (cyclic) {
        qm &= (P - 1);
        qp &= (P - 1);
      }
      if (qm >= 0) dm = ld.get(sl.at(qm));
      if (qp < P) dp = ld.get(sl.at(qp));
      const int ts = step * P + t * S + seg;
      const double k1 = st[T.k1 + ts], k2 = st[T.k2 + ts];
      nd[t] = ld.get(sl.at(q)) - k1 * dm - k2 * dp;
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < 2; t++) ld.set(sl.at(q0 + t), nd[t]);
    __syncthreads();
  }
  V xs[2];
#pragma unroll
  for (int t = 0; t < 2; t++) {
    const int q = q0 + t;
    const int ts = t * S + seg;
    if (cyclic) {
      xs[t] = st[T.fin0 + ts] * (st[T.fin1 + ts] * ld.get(sl.at(q)) -
                                 st[T.fin2 + ts] * ld.get(sl.at((q + P / 2) & (P - 1))));
    } else {
      xs[t] = st[T.fin0 + ts] * ld.get(sl.at(q));
    }
  }
  x0 = xs[0];
  x1 = xs[1];
}

// Lines the fast kernel leaves to the general one: `count` entries of
// `list`, each the first line la + nla_pad * lb of a group of adjacent
// lines (one general-kernel block).  `next` is the counter of the next
// queued sweep on the stream, which this sweep's general kernel clears.
struct LineQueue {
  unsigned* count;
  unsigned* next;
  unsigned* list;
};

// A batch of lines.  Sys provides
//   using V, Pt;  static constexpr bool kCyclic, kHasRef, kNb;
//   int n() const;                              rows per line
//   bool valid(int la, int lb) const;           line exists
//   V rhs(la, lb, r) const;                     right-hand side of row r
//   Pt pt(la, lb, r) const;                     the per-row inputs of row r's
//                                               coefficients (r = -1 and n
//                                               too when kNb: the system's
//                                               edge rule applies)
//   bool pt_ref(const Pt&) const;               inputs equal the reference
//                                               line's (cellvalue 0, ...)
//   void coef(la, lb, r, p, pm, pp, a, b, c);   row r from its inputs and, when
//                                               kNb, its neighbours'
//   void coef_ref(r, a, b, c) const;            row r of the reference line
//   void store(la, lb, r, V x) const;           solution of row r
//   const unsigned* lflag; unsigned lbit;       when lflag is set, line
//   int lstride;                                (la, lb) has a point whose
//                                               pt_ref is false only if bit
//                                               lbit of lflag[la + lstride lb]
//                                               is set
// Block (bx, by) holds lines la = bx * LPB + (line within block), lb = by.
// Every load of a thread is issued before any value is examined (a thread
// past the last line reads the block's first line, and rows past a short
// segment repeat its last row).
template <int S, int NT, bool LANE_LINES>
struct LineMap {
  static constexpr int LPB = NT / S;
  static constexpr int P = 2 * S;
  int line, seg;
  __device__ LineMap(int tid)
      : line(LANE_LINES ? (tid % LPB) : (tid / S)), seg(LANE_LINES ? (tid / LPB) : (tid % S)) {}
};

// XCD-aware linear block number: the dispatcher deals workgroups to the
// eight XCDs round-robin, so consecutive blocks (which share the 128-byte
// lines of a row when a block covers half of one) are renumbered to run
// on the same XCD at about the same time.
__device__ inline unsigned xcd_linear(unsigned p, unsigned nb) {
  const unsigned X = 8;
  const unsigned per = nb / X, rem = nb % X;
  const unsigned x = p % X, r = p / X;
  return x * per + (x < rem ? x : rem) + r;
}

// General path for one block: elimination in registers, PCR in LDS.  Each
// thread's line is (la, lb), a valid line; `ok` says whether its solution
// is stored (threads past the last line load a valid line and store
// nothing).
//
// With ONLY_DIRTY the lines whose rows all have the reference inputs (the
// ones the recorded-multiplier path has stored) are solved but not
// stored: the test is the one k_line_fast makes (the step's dirty-line
// mask when the system has one, else every row's inputs).  `lcl` holds
// one int per line of the block.
template <int S, int MMAX, int NT, bool LANE_LINES, bool ONLY_DIRTY, typename Sys>
__device__ inline void line_block_general(const Sys& sys, int la, int lb, bool ok, double* s_a,
                                          double* s_b, double* s_c, double* s_d, int* lcl) {
  using V = typename Sys::V;
  using Pt = typename Sys::Pt;
  using M = LineMap<S, NT, LANE_LINES>;
  constexpr int P = M::P;
  // the thread's segment is made opaque here so that the row coefficients'
  // metric terms, the same in every block, are not hoisted out of the
  // caller's block loop and kept in registers across it
  int tid = threadIdx.x;
  asm volatile("" : "+v"(tid));
  const M mp(tid);
  const int n = sys.n();
  const int lo = (mp.seg * n) / S;
  const int m = ((mp.seg + 1) * n) / S - lo;

  // rows' right-hand sides and coefficients; the coefficient inputs roll
  // through a three-row window (previous, current, next row)
  V d[MMAX];
  double a[MMAX], b[MMAX], c[MMAX];
  if (ONLY_DIRTY && mp.seg == 0) lcl[mp.line] = 1;
  int clean = 1;
  Pt prev = Sys::kNb ? sys.pt(la, lb, lo - 1) : Pt();
  Pt cur = sys.pt(la, lb, lo);
  if (ONLY_DIRTY && Sys::kNb) clean &= int(sys.pt_ref(prev));
#pragma unroll
  for (int r = 0; r < MMAX; r++) {
    const int q = lo + (r < m ? r : m - 1);
    d[r] = sys.rhs(la, lb, q);
    Pt next = Pt();
    if (Sys::kNb) next = sys.pt(la, lb, r < m - 1 ? q + 1 : lo + m);
    sys.coef(la, lb, q, cur, prev, next, a[r], b[r], c[r]);
    if (ONLY_DIRTY) {
      clean &= int(sys.pt_ref(cur));
      if (Sys::kNb && r == m - 1) clean &= int(sys.pt_ref(next));
    }
    if (r + 1 < MMAX) {
      prev = cur;
      cur = Sys::kNb ? next : sys.pt(la, lb, lo + (r + 1 < m ? r + 1 : m - 1));
    }
  }
  if (ONLY_DIRTY && sys.lflag) clean = int((sys.lflag[la + sys.lstride * lb] & sys.lbit) == 0u);
  TabLayout none{};
  segment_eliminate<MMAX, false, V>(lo, m, mp.seg, a, b, c, d, nullptr, none);
  const LSlot<S> sl = LANE_LINES ? LSlot<S>{mp.line, M::LPB} : LSlot<S>{mp.line * P, 1};
  LdsVec<V> ld{s_d, M::LPB * P};
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
  const int i0 = sl.at(2 * mp.seg), i1 = sl.at(2 * mp.seg + 1);
  s_a[i0] = a[0];
  s_b[i0] = 1.0;
  s_c[i0] = c[0];
  ld.set(i0, d[0]);
  s_a[i1] = alast;
  s_b[i1] = 1.0;
  s_c[i1] = clast;
  ld.set(i1, dlast);
  __syncthreads();
  if (ONLY_DIRTY && !clean) lcl[mp.line] = 0;
  V xf, xl;
  pcr_pair<S, false, V>(s_a, s_b, s_c, ld, sl, mp.seg, Sys::kCyclic, xf, xl, nullptr, none);
  if (ONLY_DIRTY) {
    __syncthreads();
    ok = ok && lcl[mp.line] == 0;
  }
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

// The general solve of every line, block (bx, by) holding lines
// la = bx * LPB + (line within block), lb = by of an nbx x nby grid of
// blocks, in XCD-aware order.
// Systems with a block-0 task (kBlock0, run when the system's `block0` is
// set): the launch adds kBlock0Pad blocks in front, block 0 runs the task
// with the block's LDS arrays as scratch and the other added blocks return,
// so the line blocks keep their XCD placement.
template <typename S, typename = void>
struct block0_task : std::false_type {};
template <typename S>
struct block0_task<S, std::void_t<decltype(S::kBlock0)>> : std::bool_constant<S::kBlock0> {};
constexpr unsigned kBlock0Pad = 8;
template <typename Sys>
inline unsigned block0_pad(const Sys& sys) {
  if constexpr (block0_task<Sys>::value) return sys.block0 ? kBlock0Pad : 0u;
  return 0u;
}

template <int S, int MMAX, int NT, bool LANE_LINES, typename Sys>
__device__ inline void line_general_body(const Sys& sys, int nbx) {
  using V = typename Sys::V;
  using M = LineMap<S, NT, LANE_LINES>;
  constexpr int LPB = NT / S;
  constexpr int P = 2 * S;
  __shared__ double s_a[LPB * P], s_b[LPB * P], s_c[LPB * P];
  __shared__ double s_d[LPB * P * vwords<V>()];
  unsigned bid = blockIdx.x, nblk = gridDim.x;
  if constexpr (block0_task<Sys>::value) {
    if (sys.block0) {
      if (bid < kBlock0Pad) {
        if (bid == 0) sys.run_block0(s_a, LPB * P, s_d, LPB * P * vwords<V>());
        return;
      }
      bid -= kBlock0Pad;
      nblk -= kBlock0Pad;
    }
  }
  const unsigned blk = xcd_linear(bid, nblk);
  const int bx = (int)(blk % (unsigned)nbx), by = (int)(blk / (unsigned)nbx);
  const M mp(threadIdx.x);
  const int la = bx * LPB + mp.line;
  const bool ok = sys.valid(la, by);
  line_block_general<S, MMAX, NT, LANE_LINES, false, Sys>(sys, ok ? la : bx * LPB, by, ok, s_a,
                                                          s_b, s_c, s_d, nullptr);
}
template <int S, int MMAX, int NT, bool LANE_LINES, typename Sys>
__global__ __launch_bounds__(NT) void k_line_general(Sys sys, int nbx, int nby) {
  (void)nby;
  line_general_body<S, MMAX, NT, LANE_LINES>(sys, nbx);
}
// The same with the registers held to 128 per lane (4 waves per SIMD), for
// the systems that declare kWaves4: the spectral solve fits in 128 with no
// spill and runs 8% faster than at its natural 130 (3 waves per SIMD).
template <int S, int MMAX, int NT, bool LANE_LINES, typename Sys>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4))) void k_line_general_w4(
    Sys sys, int nbx, int nby) {
  (void)nby;
  line_general_body<S, MMAX, NT, LANE_LINES>(sys, nbx);
}
template <typename S, typename = void>
struct waves4 : std::false_type {};
template <typename S>
struct waves4<S, std::void_t<decltype(S::kWaves4)>> : std::bool_constant<S::kWaves4> {};

// The general solve of the line groups the fast kernel queued: each entry
// is the first line (la0, lb) of LPB adjacent lines, which a block loads as
// the fast kernel does (coalesced) and of which it stores the lines the
// fast kernel left.  The grid strides over the queue; block 0 clears the
// next sweep's counter (the previous sweep's general kernel, which read it,
// has finished, and the next sweep's fast kernel has not started).
template <int S, int MMAX, int NT, bool LANE_LINES, typename Sys>
__global__ __launch_bounds__(NT) void k_line_general_q(Sys sys, LineQueue Q, unsigned nla_pad) {
  using V = typename Sys::V;
  using M = LineMap<S, NT, LANE_LINES>;
  constexpr int LPB = NT / S;
  constexpr int P = 2 * S;
  __shared__ double s_a[LPB * P], s_b[LPB * P], s_c[LPB * P];
  __shared__ double s_d[LPB * P * vwords<V>()];
  __shared__ int lcl[LPB];
  const unsigned count = __atomic_load_n(Q.count, __ATOMIC_RELAXED);
  if (blockIdx.x == 0 && threadIdx.x == 0) *Q.next = 0u;
  const M mp(threadIdx.x);
  for (unsigned gi = blockIdx.x; gi < count; gi += gridDim.x) {
    const unsigned id = Q.list[gi];
    const int la0 = (int)(id % nla_pad), lb = (int)(id / nla_pad);
    const int la = la0 + mp.line;
    const bool ok = sys.valid(la, lb);
    line_block_general<S, MMAX, NT, LANE_LINES, true, Sys>(sys, ok ? la : la0, lb, ok, s_a, s_b,
                                                           s_c, s_d, lcl);
    __syncthreads();
  }
}

// The recorded-multiplier path.  The reference line's table (validated for
// this call by k_line_factor, which also keeps a copy in the LDS order of
// LdsTab, so that staging it is one contiguous copy) is staged in LDS once
// per block and together with its first data loads; the block then takes GPB
// groups of LPB lines in turn.  A line whose rows all have the reference
// inputs applies the recorded multipliers: its rows then have the
// reference coefficients bit for bit, so it performs the same operations on
// d as the general path.  Any other line is queued for k_line_general_q and
// not written here.
template <int S, int MMAX, int NT, bool LANE_LINES, int LPBG, typename Sys>
__global__ __launch_bounds__(NT) void k_line_fast(Sys sys, const double* __restrict__ tab,
                                                  TabLayout L, LineQueue Q, int nbx,
                                                  unsigned ngrp) {
  using V = typename Sys::V;
  using M = LineMap<S, NT, LANE_LINES>;
  constexpr int P = M::P;
  extern __shared__ double smem[];
  __shared__ int lclean[M::LPB];
  const LdsTab<S, MMAX> T(L.steps);
  double* st = smem;
  double* sd = smem + T.total;

  // a block takes GPB consecutive groups of lines and stages the table once
  // (x lines: four lines per block would otherwise stage a table larger
  // than their data; y and z blocks hold 16-32 lines and take one group)
  constexpr int GPB = LANE_LINES ? 1 : 8;
  const unsigned first = xcd_linear(blockIdx.x, gridDim.x) * (unsigned)GPB;
  const M mp(threadIdx.x);
  const int seg = mp.seg;
  const int n = sys.n();
  const int lo = (seg * n) / S;
  const int m = ((seg + 1) * n) / S - lo;
#pragma unroll 1
  for (int gi = 0; gi < GPB; gi++) {
    const unsigned blk = first + (unsigned)gi;
    if (blk >= ngrp) break;  // the same for every thread of the block
    // the previous group's reads of lclean and of the reduced system are done
    if (gi > 0) __syncthreads();
    const int bx = (int)(blk % (unsigned)nbx), by = (int)(blk / (unsigned)nbx);
    const int la = bx * M::LPB + mp.line, lb = by;
    const bool ok = sys.valid(la, lb);
    const int lc = ok ? la : bx * M::LPB;
    if (seg == 0) lclean[mp.line] = 1;

    V d[MMAX];
    int clean = 1;
#pragma unroll
    for (int r = 0; r < MMAX; r++) {
      const int q = lo + (r < m ? r : m - 1);
      d[r] = sys.rhs(lc, lb, q);
    }
    if (sys.lflag) {
      // the step's dirty-line mask stands for the per-point test
      clean = int((sys.lflag[lc + sys.lstride * lb] & sys.lbit) == 0u);
    } else {
#pragma unroll
      for (int r = 0; r < MMAX; r++) {
        const int q = lo + (r < m ? r : m - 1);
        clean &= int(sys.pt_ref(sys.pt(lc, lb, q)));
      }
      if (Sys::kNb) clean &= int(sys.pt_ref(sys.pt(lc, lb, lo - 1))) &
                             int(sys.pt_ref(sys.pt(lc, lb, lo + m)));
    }
    if (gi == 0) {
      // stage the table while the first group's rows are in flight: KT
      // loads in flight per thread before the first wait
      constexpr int KT = 8;
      for (int u0 = 0; u0 < T.total; u0 += KT * NT) {
        double v[KT];
#pragma unroll
        for (int k = 0; k < KT; k++) {
          const int u = u0 + k * NT + (int)threadIdx.x;
          v[k] = tab[u < T.total ? u : T.total - 1];
        }
#pragma unroll
        for (int k = 0; k < KT; k++) {
          const int u = u0 + k * NT + (int)threadIdx.x;
          if (u < T.total) st[u] = v[k];
        }
      }
    }
    __syncthreads();
    if (!clean) lclean[mp.line] = 0;
    // the segment elimination's operations on d, with the recorded multipliers
#pragma unroll
    for (int r = 0; r < MMAX; r++) {
      if (r < m) {
        const double f = st[T.f + r * S + seg];
        if (r < 2)
          d[r] = f * d[r];
        else
          d[r] = f * (d[r] - st[T.g + r * S + seg] * d[r - 1]);
      }
    }
#pragma unroll
    for (int r = MMAX - 3; r >= 1; r--)
      if (r <= m - 3) d[r] = d[r] - st[T.h + r * S + seg] * d[r + 1];
    if (m >= 3) d[0] = st[T.f0 + seg] * (d[0] - st[T.h + seg] * d[1]);
    V dlast;
    vzero(dlast);
#pragma unroll
    for (int r = 0; r < MMAX; r++)
      if (r == m - 1) dlast = d[r];
    const LSlot<S> sl = LANE_LINES ? LSlot<S>{mp.line, M::LPB} : LSlot<S>{mp.line * P, 1};
    LdsVec<V> ld{sd, M::LPB * P};
    ld.set(sl.at(2 * seg), d[0]);
    ld.set(sl.at(2 * seg + 1), dlast);
    __syncthreads();
    const bool line_clean = lclean[mp.line] != 0;
    // queue each group of LPBG adjacent lines (the general kernel's block)
    // that holds a valid line left unsolved here
    static_assert(M::LPB % LPBG == 0, "general groups must tile the fast block");
    if ((int)threadIdx.x < M::LPB / LPBG) {
      const int l0 = (int)threadIdx.x * LPBG, la0 = bx * M::LPB + l0;
      int dirty = 0;
#pragma unroll
      for (int q = 0; q < LPBG; q++) dirty |= int(lclean[l0 + q] == 0 && sys.valid(la0 + q, lb));
      if (dirty)
        Q.list[atomicAdd(Q.count, 1u)] = (unsigned)la0 + (unsigned)(nbx * M::LPB) * (unsigned)lb;
    }
    V xf, xl;
    pcr_apply<S, MMAX, V>(ld, sl, seg, Sys::kCyclic, xf, xl, st, T);
    if (ok && line_clean) {
#pragma unroll
      for (int r = 0; r < MMAX; r++) {
        if (r < m) {
          V x;
          if (r == 0)
            x = xf;
          else if (r == m - 1)
            x = xl;
          else
            x = d[r] - st[T.afin + r * S + seg] * xf - st[T.cfin + r * S + seg] * xl;
          sys.store(la, lb, lo + r, x);
        }
      }
    }
  }
}

// Validates the reference line's table for this call and refactors it when
// needed (one block, one thread per segment): the table is kept when its
// recorded coefficients equal the reference line's coefficients now and
// `fresh` is false; otherwise the line is factored again and every
// multiplier applied to the right-hand side is recorded.
template <int S, int MMAX, typename Sys>
__global__ __launch_bounds__(S) void k_line_factor(Sys sys, double* tab, TabLayout L, int fresh,
                                                   double* tabs) {
  constexpr int P = 2 * S;
  __shared__ double s_a[P], s_b[P], s_c[P], s_d[P];
  const int seg = threadIdx.x;
  const int n = sys.n();
  const int lo = (seg * n) / S;
  const int m = ((seg + 1) * n) / S - lo;
  double a[MMAX], b[MMAX], c[MMAX], d[MMAX];
  int same = fresh ? 0 : 1;
  // rows past a short segment repeat its last row, so that every load is
  // unconditional and all are in flight together
#pragma unroll
  for (int r = 0; r < MMAX; r++) {
    const int q = lo + (r < m ? r : m - 1);
    sys.coef_ref(q, a[r], b[r], c[r]);
    d[r] = 0.0;
    same &= int(a[r] == tab[L.o_refa + q]) & int(b[r] == tab[L.o_refb + q]) &
            int(c[r] == tab[L.o_refc + q]);
  }
  if (__syncthreads_and(same)) return;
#pragma unroll
  for (int r = 0; r < MMAX; r++)
    if (r < m) {
      tab[L.o_refa + lo + r] = a[r];
      tab[L.o_refb + lo + r] = b[r];
      tab[L.o_refc + lo + r] = c[r];
    }
  segment_eliminate<MMAX, true, double>(lo, m, seg, a, b, c, d, tab, L);
  double alast = 0.0, clast = 0.0;
#pragma unroll
  for (int r = 0; r < MMAX; r++)
    if (r == m - 1) {
      alast = a[r];
      clast = c[r];
    }
  const LSlot<S> sl{0, 1};
  const int i0 = sl.at(2 * seg), i1 = sl.at(2 * seg + 1);
  s_a[i0] = a[0];
  s_b[i0] = 1.0;
  s_c[i0] = c[0];
  s_d[i0] = 0.0;
  s_a[i1] = alast;
  s_b[i1] = 1.0;
  s_c[i1] = clast;
  s_d[i1] = 0.0;
  __syncthreads();
  double x0, x1;
  pcr_pair<S, true, double>(s_a, s_b, s_c, LdsVec<double>{s_d, P}, sl, seg, Sys::kCyclic, x0, x1,
                            tab, L);
  // the copy in LDS order (LdsTab::src), read back after every thread's
  // table stores are visible to the block
  __threadfence();
  __syncthreads();
  const LdsTab<S, MMAX> T(L.steps);
  for (int u = seg; u < T.total; u += S) tabs[u] = tab[T.src(u, n, L)];
}

// The queue shared by the sweeps on the library stream, sized for at
// least `nlines` entries.
LineQueue line_queue(unsigned nlines, hipStream_t s);

// Host cache of reference-line tables, keyed by the system's description;
// *validated is set when the table was validated earlier in the same
// entry-point call (whose metric arrays and scalars do not change).
double* line_table(const void* key, size_t keylen, long total, bool* fresh, bool* validated);

// One sweep: the general kernel uses SG segments of at most MG rows; the
// fast kernel SF segments of at most MF rows (fewer, longer segments: its
// rows need one register each, and fewer segments mean a smaller staged
// table and fewer PCR steps).  Both use NT-thread blocks.
template <bool LANE_LINES, int SG, int MG, int SF, int MF, int NT, typename Sys, int NTG = NT>
inline int line_sweep(const Sys& sys, int n, int nla, int nlb, hipStream_t s) {
  constexpr int LPB_G = NTG / SG, LPB_F = NT / SF;
  double* tab = nullptr;
  bool fresh = false, validated = false;
  TabLayout L = tab_layout(n, SF, Sys::kCyclic);
  const size_t lds = sizeof(double) * ((size_t)LdsTab<SF, MF>(L.steps).total +
                                       (size_t)LPB_F * 2 * SF * vwords<typename Sys::V>());
  const long staged = LdsTab<SF, MF>(L.steps).total;
  if (Sys::kHasRef && lds <= 65536 && (n + SF - 1) / SF <= MF && n >= 2 * SF) {
    auto key = sys.ref_key(n, SF);
    // the table and its copy in LDS order; null: general only
    tab = line_table(&key, sizeof(key), L.total + staged, &fresh, &validated);
  }
  if (tab) {
    const int nbx = (nla + LPB_F - 1) / LPB_F;
    const unsigned nb = (unsigned)nbx * (unsigned)nlb;
    const unsigned nla_pad = (unsigned)nbx * (unsigned)LPB_F;
    LineQueue Q = line_queue(nla_pad * (unsigned)nlb, s);
    if (!Q.list) return -2;
    constexpr unsigned GPB = LANE_LINES ? 1u : 8u;  // groups per block, as in k_line_fast
    const unsigned ngb = (nb + GPB - 1) / GPB;
    double* tabs = tab + L.total;
    if (!validated)
      k_line_factor<SF, MF, Sys><<<1, SF, 0, s>>>(sys, tab, L, fresh ? 1 : 0, tabs);
    k_line_fast<SF, MF, NT, LANE_LINES, LPB_G, Sys><<<ngb, NT, lds, s>>>(sys, tabs, L, Q, nbx, nb);
    const unsigned ngen = nb < 1024u ? nb : 1024u;
    k_line_general_q<SG, MG, NTG, LANE_LINES, Sys><<<ngen, NTG, 0, s>>>(sys, Q, nla_pad);
  } else {
    const int nbx = (nla + LPB_G - 1) / LPB_G;
    const unsigned nb = (unsigned)nbx * (unsigned)nlb + block0_pad(sys);
    if constexpr (waves4<Sys>::value && MG <= 8)  // longer segments would spill
      k_line_general_w4<SG, MG, NTG, LANE_LINES, Sys><<<nb, NTG, 0, s>>>(sys, nbx, nlb);
    else
      k_line_general<SG, MG, NTG, LANE_LINES, Sys><<<nb, NTG, 0, s>>>(sys, nbx, nlb);
  }
  return 0;
}

// Launch helper: the general kernel's segment count SG is the smallest
// power of two (at most 64) that leaves at most 8 rows per segment (12 or
// 16 when 64 segments are not enough), so its reduced system has 2 SG rows;
// the fast kernel takes half as many segments, of at most 12 or 16 rows.  A
// sweep with a reference table is three launches: validate the table, the
// fast kernel over every item, the general kernel over the queued ones.
template <bool LANE_LINES, typename Sys>
inline int launch_line_solve(const Sys& sys, int n, int nla, int nlb, hipStream_t s) {
  int S = 1;
  while (S < 64 && (n + S - 1) / S > 8) S <<= 1;
  const int mmax = (n + S - 1) / S;
  if (n < 2 || mmax > 16 || n < 2 * S) return -1;
  const bool f12 = S > 1 && (n + S / 2 - 1) / (S / 2) <= 12;
  if constexpr (!Sys::kHasRef && LANE_LINES) {
    // systems with no reference line (the spectral solve, the transposed
    // sweeps) run the general kernel alone, whose registers allow 3 waves
    // per SIMD: 4-wave blocks fill the CUs better than 8-wave ones
    if (mmax <= 8) {
      switch (S) {
        case 64: return line_sweep<LANE_LINES, 64, 8, 64, 8, 256>(sys, n, nla, nlb, s);
        case 32: return line_sweep<LANE_LINES, 32, 8, 32, 8, 256>(sys, n, nla, nlb, s);
        case 16: return line_sweep<LANE_LINES, 16, 8, 16, 8, 256>(sys, n, nla, nlb, s);
        case 8: return line_sweep<LANE_LINES, 8, 8, 8, 8, 256>(sys, n, nla, nlb, s);
        case 4: return line_sweep<LANE_LINES, 4, 8, 4, 8, 256>(sys, n, nla, nlb, s);
        case 2: return line_sweep<LANE_LINES, 2, 8, 2, 8, 256>(sys, n, nla, nlb, s);
        default: return line_sweep<LANE_LINES, 1, 8, 1, 8, 256>(sys, n, nla, nlb, s);
      }
    }
  }
  // the general kernels hold 116-200 VGPRs (2-4 waves per SIMD): 4-wave
  // blocks for them, 8-wave blocks for the fast kernels
#define VK_SW(SG, MG, SF, MF) \
  return line_sweep<LANE_LINES, SG, MG, SF, MF, 512, Sys, 256>(sys, n, nla, nlb, s)
  if (mmax > 8) {
    // long lines: 64 segments of up to 12 rows keep the fast kernel's
    // staged table within 64 KB of LDS (768-row lines), 16 rows beyond
    if ((n + 63) / 64 <= 12) VK_SW(64, 12, 64, 12);
    VK_SW(64, 16, 64, 16);
  }
  if (!LANE_LINES) {
    // lines contiguous in memory: a thread's rows are consecutive words, and
    // short segments keep the lanes' loads close together
    switch (S) {
      case 64: return line_sweep<LANE_LINES, 64, 8, 64, 8, 256>(sys, n, nla, nlb, s);
      case 32: return line_sweep<LANE_LINES, 32, 8, 32, 8, 256>(sys, n, nla, nlb, s);
      case 16: return line_sweep<LANE_LINES, 16, 8, 16, 8, 256>(sys, n, nla, nlb, s);
      case 8: return line_sweep<LANE_LINES, 8, 8, 8, 8, 256>(sys, n, nla, nlb, s);
      case 4: return line_sweep<LANE_LINES, 4, 8, 4, 8, 256>(sys, n, nla, nlb, s);
      case 2: return line_sweep<LANE_LINES, 2, 8, 2, 8, 256>(sys, n, nla, nlb, s);
      default: return line_sweep<LANE_LINES, 1, 8, 1, 8, 256>(sys, n, nla, nlb, s);
    }
  }
  switch (S) {
    case 64: if (f12) VK_SW(64, 8, 32, 12); VK_SW(64, 8, 32, 16);
    case 32: if (f12) VK_SW(32, 8, 16, 12); VK_SW(32, 8, 16, 16);
    case 16: if (f12) VK_SW(16, 8, 8, 12); VK_SW(16, 8, 8, 16);
    case 8: if (f12) VK_SW(8, 8, 4, 12); VK_SW(8, 8, 4, 16);
    case 4: if (f12) VK_SW(4, 8, 2, 12); VK_SW(4, 8, 2, 16);
    case 2: if (f12) VK_SW(2, 8, 1, 12); VK_SW(2, 8, 1, 16);
    default: VK_SW(1, 8, 1, 8);
  }
#undef VK_SW
}

}  // namespace vk
