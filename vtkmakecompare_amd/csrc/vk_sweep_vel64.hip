// vk_sweep_vel64.hip -- instantiation of launch_adi_velocity<double> (vk_linesys.h).
#include "vk_linesys.h"

namespace vk {

template int launch_velocity_sweep<double>(const Geom&, const Metrics&, const RMetrics&, double*,
                                           const double*, int, int, double, double, hipStream_t,
                                           const LineFlags*);
template int launch_adi_velocity<double>(const Geom&, const Metrics&, const RMetrics&, double*,
                                         const double*, int, double, double, hipStream_t,
                                         const LineFlags*);

}  // namespace vk
