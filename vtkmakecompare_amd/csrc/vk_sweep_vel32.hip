// vk_sweep_vel32.hip -- instantiation of launch_adi_velocity<float> (vk_linesys.h).
#include "vk_linesys.h"

namespace vk {

template int launch_velocity_sweep<float>(const Geom&, const Metrics&, const RMetrics&, double*,
                                          const float*, int, int, double, double, hipStream_t,
                                          const LineFlags*);
template int launch_adi_velocity<float>(const Geom&, const Metrics&, const RMetrics&, double*,
                                        const float*, int, double, double, hipStream_t,
                                        const LineFlags*);

}  // namespace vk
