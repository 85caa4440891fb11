// vk_adj_gather.hip -- the explicit phase's transposed stencils: for every
// input cell, the gather over the outputs that read it, each Jacobian entry
// the forward stencil template evaluated on a dual number seeded at that
// cell.
//
// Built with -ffinite-math-only -fno-signed-zeros on top of the library's
// -ffp-contract=off.  A seeded read's derivative is a compile-time 1 or 0
// (the seed test folds to a comparison of constant offsets), and with
// these flags the compiler may fold 0 * x to 0 and x + 0 to x, so only the
// products on the seed's path are evaluated and every value part that
// feeds no derivative is dropped.  No reassociation or contraction is
// allowed, so each surviving term is rounded as before; a term the fold
// removes was an exact zero, and the sums it joined are unchanged except
// for the sign of a zero.
#include "vk_adjoint.h"

namespace vk {

// Where each explicit term at p reads field F (0, 1, 2, 3 = u, v, w, T):
// the offsets q - p of its reads, from the expressions of vk_stencil_ops.h.
// Terms: 0 conv_u_r, 1 conv_v_r, 2 conv_w_r, 3 cond of the field's own
// component, 4 conv_scalar_r, 5 diff_scalar_r.  A term that does not read
// F at q - p has zero derivative there and is not evaluated.
constexpr bool seven_point(int a, int b, int c) {
  return (a < 0 ? -a : a) + (b < 0 ? -b : b) + (c < 0 ? -c : c) <= 1;
}
constexpr bool is_one_of(int a, int b, int c, const int (&o)[4][3]) {
  for (int q = 0; q < 4; q++)
    if (o[q][0] == a && o[q][1] == b && o[q][2] == c) return true;
  return false;
}
template <int F, int TERM>
constexpr bool term_reads(int a, int b, int c) {
  // the cross terms' four-point footprints
  constexpr int u_in_v[4][3] = {{0, 1, 0}, {0, 0, 0}, {-1, 1, 0}, {-1, 0, 0}};
  constexpr int u_in_w[4][3] = {{0, 0, 1}, {0, 0, 0}, {-1, 0, 1}, {-1, 0, 0}};
  constexpr int v_in_u[4][3] = {{1, 0, 0}, {0, 0, 0}, {1, -1, 0}, {0, -1, 0}};
  constexpr int v_in_w[4][3] = {{0, 0, 1}, {0, 0, 0}, {0, -1, 1}, {0, -1, 0}};
  constexpr int w_in_u[4][3] = {{1, 0, 0}, {0, 0, 0}, {1, 0, -1}, {0, 0, -1}};
  constexpr int w_in_v[4][3] = {{0, 1, 0}, {0, 0, 0}, {0, 1, -1}, {0, 0, -1}};
  if (F == 0) {
    if (TERM == 0 || TERM == 3) return seven_point(a, b, c);
    if (TERM == 1) return is_one_of(a, b, c, u_in_v);
    if (TERM == 2) return is_one_of(a, b, c, u_in_w);
    if (TERM == 4) return b == 0 && c == 0 && (a == 0 || a == -1);
    return false;
  }
  if (F == 1) {
    if (TERM == 1 || TERM == 3) return seven_point(a, b, c);
    if (TERM == 0) return is_one_of(a, b, c, v_in_u);
    if (TERM == 2) return is_one_of(a, b, c, v_in_w);
    if (TERM == 4) return a == 0 && c == 0 && (b == 0 || b == -1);
    return false;
  }
  if (F == 2) {
    if (TERM == 2 || TERM == 3) return seven_point(a, b, c);
    if (TERM == 0) return is_one_of(a, b, c, w_in_u);
    if (TERM == 1) return is_one_of(a, b, c, w_in_v);
    if (TERM == 4) return a == 0 && b == 0 && (c == 0 || c == -1);
    return false;
  }
  return (TERM == 4 || TERM == 5) && seven_point(a, b, c);
}

// The contribution to lam(q) of output point p = q - (DI, DJ, DK): each
// term that reads F at q, evaluated on duals seeded at q and weighted by
// the term's adjoint at p.  Terms are added in the order uv, vv, wv, the
// conduction, vt (dtt for T), as a loop over all terms would add them.
template <int F, int DI, int DJ, int DK>
__device__ inline void gather_from(const Geom& g, const Metrics& m, const RMetrics& r,
                                   const StepIn& in, const Weights& wt, const Dd& ddt,
                                   const Dd& dc0, const Dd& dcd, int qi, int qj, int qk,
                                   double& acc) {
  constexpr bool t0 = term_reads<F, 0>(DI, DJ, DK), t1 = term_reads<F, 1>(DI, DJ, DK),
                 t2 = term_reads<F, 2>(DI, DJ, DK), t3 = term_reads<F, 3>(DI, DJ, DK),
                 t4 = term_reads<F, 4>(DI, DJ, DK), t5 = term_reads<F, 5>(DI, DJ, DK);
  if constexpr (!(t0 || t1 || t2 || t3 || t4 || t5)) {
    return;
  } else {
    const int i = qi - DI, j = qj - DJ, k = qk - DK;
    if (!is_interior(g, i, j, k)) return;
    const long long p = g.at(i, j, k);
    const PlainV RC{in.rc, g}, K{in.thc, g};
    const Dd rrc(1.0 / in.rc[p]);
    if constexpr (F == 0) {
      const SeedV U{in.u, g, qi, qj, qk};
      const PlainV V{in.v, g}, W{in.w, g}, T{in.t, g};
      if constexpr (t0) acc += wt.uv[p] * conv_u_r<Dd>(U, V, W, m, r, i, j, k, ddt).d;
      if constexpr (t1) acc += wt.vv[p] * conv_v_r<Dd>(U, V, W, m, r, i, j, k, ddt).d;
      if constexpr (t2) acc += wt.wv[p] * conv_w_r<Dd>(U, V, W, m, r, i, j, k, ddt).d;
      if constexpr (t3) acc += wt.ud[p] * cond_u_r<Dd>(U, r, i, j, k, dc0).d;
      if constexpr (t4)
        acc += wt.vt[p] * conv_scalar_r<Dd>(U, V, W, T, RC, m, r, i, j, k, ddt, rrc).d;
    } else if constexpr (F == 1) {
      const SeedV V{in.v, g, qi, qj, qk};
      const PlainV U{in.u, g}, W{in.w, g}, T{in.t, g};
      if constexpr (t0) acc += wt.uv[p] * conv_u_r<Dd>(U, V, W, m, r, i, j, k, ddt).d;
      if constexpr (t1) acc += wt.vv[p] * conv_v_r<Dd>(U, V, W, m, r, i, j, k, ddt).d;
      if constexpr (t2) acc += wt.wv[p] * conv_w_r<Dd>(U, V, W, m, r, i, j, k, ddt).d;
      if constexpr (t3) acc += wt.vd[p] * cond_v_r<Dd>(V, r, i, j, k, dc0).d;
      if constexpr (t4)
        acc += wt.vt[p] * conv_scalar_r<Dd>(U, V, W, T, RC, m, r, i, j, k, ddt, rrc).d;
    } else if constexpr (F == 2) {
      const SeedV W{in.w, g, qi, qj, qk};
      const PlainV U{in.u, g}, V{in.v, g}, T{in.t, g};
      if constexpr (t0) acc += wt.uv[p] * conv_u_r<Dd>(U, V, W, m, r, i, j, k, ddt).d;
      if constexpr (t1) acc += wt.vv[p] * conv_v_r<Dd>(U, V, W, m, r, i, j, k, ddt).d;
      if constexpr (t2) acc += wt.wv[p] * conv_w_r<Dd>(U, V, W, m, r, i, j, k, ddt).d;
      if constexpr (t3) acc += wt.wd[p] * cond_w_r<Dd>(W, r, i, j, k, dc0).d;
      if constexpr (t4)
        acc += wt.vt[p] * conv_scalar_r<Dd>(U, V, W, T, RC, m, r, i, j, k, ddt, rrc).d;
    } else {
      const SeedV T{in.t, g, qi, qj, qk};
      const PlainV U{in.u, g}, V{in.v, g}, W{in.w, g};
      if constexpr (t4)
        acc += wt.vt[p] * conv_scalar_r<Dd>(U, V, W, T, RC, m, r, i, j, k, ddt, rrc).d;
      if constexpr (t5) acc += wt.dtt[p] * diff_scalar_r<Dd>(T, K, m, r, i, j, k, dcd, rrc).d;
    }
  }
}

// offsets in the order dk, dj, di = -1, 0, 1 (di fastest)
template <int F, int IDX>
__device__ inline void gather_all(const Geom& g, const Metrics& m, const RMetrics& r,
                                  const StepIn& in, const Weights& wt, const Dd& ddt,
                                  const Dd& dc0, const Dd& dcd, int qi, int qj, int qk,
                                  double& acc) {
  if constexpr (IDX < 27) {
    gather_from<F, IDX % 3 - 1, (IDX / 3) % 3 - 1, IDX / 9 - 1>(g, m, r, in, wt, ddt, dc0, dcd,
                                                                qi, qj, qk, acc);
    gather_all<F, IDX + 1>(g, m, r, in, wt, ddt, dc0, dcd, qi, qj, qk, acc);
  }
}

// For every cell q (halos included): the sum over interior outputs p that
// read field F at q of weight(p) * d f_p / d F(q), each derivative the
// stencil evaluated on duals seeded at q.  F = 0, 1, 2, 3 for u, v, w, T.
template <int F>
__global__ __launch_bounds__(128) void k_adj_gather(Geom g, Metrics m, RMetrics r, StepIn in,
                                                    Weights wt, double dt, double c0, double cd,
                                                    double* lam) {
  ALL_CELLS;
  const Dd ddt(dt), dc0(c0), dcd(cd);
  double acc = 0.0;
  gather_all<F, 0>(g, m, r, in, wt, ddt, dc0, dcd, qi, qj, qk, acc);
  lam[q_] += acc;
}

void launch_adj_gather(int field, const Geom& g, const Metrics& m, const RMetrics& r,
                       const StepIn& in, const Weights& wt, double dt, double c0, double cd,
                       double* lam, hipStream_t s) {
  const dim3 grid((unsigned)((g.cells() + 127) / 128));
  switch (field) {
    case 0: k_adj_gather<0><<<grid, 128, 0, s>>>(g, m, r, in, wt, dt, c0, cd, lam); break;
    case 1: k_adj_gather<1><<<grid, 128, 0, s>>>(g, m, r, in, wt, dt, c0, cd, lam); break;
    case 2: k_adj_gather<2><<<grid, 128, 0, s>>>(g, m, r, in, wt, dt, c0, cd, lam); break;
    default: k_adj_gather<3><<<grid, 128, 0, s>>>(g, m, r, in, wt, dt, c0, cd, lam); break;
  }
}

}  // namespace vk
