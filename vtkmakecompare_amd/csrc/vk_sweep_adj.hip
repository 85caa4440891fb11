// vk_sweep_adj.hip -- transposed line solves of the ADI sweeps for the
// adjoint of the FP64 time step (FP64; the recorded-multiplier path on the
// lines the forward step's masks mark clean, the general path on the rest).
#include "vk_linesys.h"

namespace vk {

int launch_velocity_sweep_adj(const Geom& g, const Metrics& m, const RMetrics& r, double* lam,
                              const double* cv, int comp, int dir, double beta, double dteta,
                              hipStream_t s, const LineFlags* fl) {
  return with_velocity_system<double>(g, m, r, lam, cv, comp, dir, beta, dteta,
                                      SolveAdj{s, lam}, fl);
}

int launch_temperature_sweep_adj(const Geom& g, const Metrics& m, const RMetrics& r,
                                 double* lam, const double* cv, const double* rc_,
                                 const double* thc, int dir, double beta, double dteta,
                                 hipStream_t s, const LineFlags* fl) {
  return with_temperature_system<double>(g, m, r, lam, cv, rc_, thc, dir, beta, dteta,
                                         SolveAdj{s, lam}, nullptr, fl);
}

}  // namespace vk
