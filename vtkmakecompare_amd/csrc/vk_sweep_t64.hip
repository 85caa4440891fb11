// vk_sweep_t64.hip -- instantiation of launch_adi_temperature<double> (vk_linesys.h).
#include "vk_linesys.h"

namespace vk {

template int launch_temperature_sweep<double>(const Geom&, const Metrics&, const RMetrics&,
                                              double*, const double*, const double*,
                                              const double*, int, double, double, hipStream_t,
                                              double*, const LineFlags*);
template int launch_adi_temperature<double>(const Geom&, const Metrics&, const RMetrics&,
                                            double*, const double*, const double*,
                                            const double*, double, double, hipStream_t,
                                            double*, const LineFlags*);

}  // namespace vk
