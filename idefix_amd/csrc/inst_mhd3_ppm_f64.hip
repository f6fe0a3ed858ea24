// Kernel instantiations for one group of configurations (generated layout, one
// translation unit per group so the groups compile in parallel).
#include "launch.hpp"

namespace idfx {
void register_mhd3_ppm_f64(Ops *t) {
  t[ops_key(3,3,1,4,2,0)] = make_ops<Cfg<3,3,true,4,2>, double>();
}
}  // namespace idfx
