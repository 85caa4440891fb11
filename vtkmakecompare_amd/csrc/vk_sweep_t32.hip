// vk_sweep_t32.hip -- instantiation of launch_adi_temperature<float> (vk_linesys.h).
#include "vk_linesys.h"

namespace vk {

template int launch_temperature_sweep<float>(const Geom&, const Metrics&, const RMetrics&,
                                             float*, const float*, const float*, const float*,
                                             int, double, double, hipStream_t, float*,
                                             const LineFlags*);
template int launch_adi_temperature<float>(const Geom&, const Metrics&, const RMetrics&, float*,
                                           const float*, const float*, const float*, double,
                                           double, hipStream_t, float*, const LineFlags*);

}  // namespace vk
