// vk_adjoint.h -- what the adjoint step's translation units share: dual
// numbers, the seeded field view, the cell indexing of the backward
// kernels, the explicit phase's inputs and weights, and the gather launcher.
#pragma once

#include "vk_kernels.h"
#include "vk_stencil_ops.h"

namespace vk {

// ------------------------------------------------------- dual numbers

// value and one directional derivative
struct Dd {
  double v, d;
  __device__ Dd() : v(0.0), d(0.0) {}
  __device__ Dd(double x) : v(x), d(0.0) {}
  __device__ Dd(double x, double y) : v(x), d(y) {}
};
__device__ inline Dd operator+(Dd a, Dd b) { return Dd(a.v + b.v, a.d + b.d); }
__device__ inline Dd operator-(Dd a, Dd b) { return Dd(a.v - b.v, a.d - b.d); }
__device__ inline Dd operator-(Dd a) { return Dd(-a.v, -a.d); }
__device__ inline Dd operator*(Dd a, Dd b) { return Dd(a.v * b.v, a.d * b.v + a.v * b.d); }
__device__ inline Dd operator/(Dd a, Dd b) {
  return Dd(a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v));
}

// a halo field read with derivative 1 at cell (si, sj, sk), 0 elsewhere
struct SeedV {
  const double* __restrict__ p;
  Geom g;
  int si, sj, sk;
  __device__ Dd operator()(int i, int j, int k) const {
    return Dd(p[g.at(i, j, k)], (i == si && j == sj && k == sk) ? 1.0 : 0.0);
  }
};
using PlainV = F3as<double, Dd>;

__device__ inline bool is_interior(const Geom& g, int i, int j, int k) {
  return i >= 1 && i <= g.nx && j >= 1 && j <= g.ny && k >= 1 && k <= g.nz;
}

#define ALL_CELLS                                                        \
  const long long q_ = (long long)blockIdx.x * blockDim.x + threadIdx.x; \
  if (q_ >= (long long)g.cells()) return;                                \
  const int qi = (int)(q_ % g.sy);                                       \
  const int qj = (int)((q_ / g.sy) % (g.ny + 2));                        \
  const int qk = (int)(q_ / g.sz)

struct Weights {      // adjoints of the explicit terms
  double *uv, *vv, *wv, *ud, *vd, *wd, *vt, *dtt;
};
struct StepIn {       // the step's inputs after SET_BC, and what the tape holds
  const double *u, *v, *w, *t, *rc, *thc, *u_ta, *v_ta, *w_ta, *cv;
  const double *su, *sv, *sw, *st;  // explicit outputs u_temp.., T work
};

// lam(q) += sum over the interior outputs p that read field F (0, 1, 2, 3 =
// u, v, w, T) at q of weight(p) * d f_p / d F(q), for every cell q
void launch_adj_gather(int field, const Geom& g, const Metrics& m, const RMetrics& r,
                       const StepIn& in, const Weights& wt, double dt, double c0, double cd,
                       double* lam, hipStream_t s);

}  // namespace vk
