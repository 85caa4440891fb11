// vk_sweep_spec.hip -- instantiations of launch_spn and launch_spn_sp0b in double and float
// (vk_linesys.h).
#include "vk_linesys.h"

namespace vk {

template int launch_spn<double>(const Geom&, double*, const double*, const double*, hipStream_t);
template int launch_spn<float>(const Geom&, float*, const double*, const double*, hipStream_t);
template int launch_spn_sp0b<double>(const Geom&, double*, const double*, const double*,
                                     hipStream_t);
template int launch_spn_sp0b<float>(const Geom&, float*, const double*, const double*,
                                    hipStream_t);

}  // namespace vk
