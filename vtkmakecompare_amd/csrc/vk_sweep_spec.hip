// vk_sweep_spec.hip -- instantiation of launch_spn<double>, launch_spn<float> (vk_linesys.h).
#include "vk_linesys.h"

namespace vk {

template int launch_spn<double>(const Geom&, double*, const double*, const double*, hipStream_t);
template int launch_spn<float>(const Geom&, float*, const double*, const double*, hipStream_t);

}  // namespace vk
