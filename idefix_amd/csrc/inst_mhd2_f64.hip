// Kernel instantiations for one group of configurations (generated layout, one
// translation unit per group so the groups compile in parallel).
#include "launch.hpp"

namespace idfx {
void register_mhd2_f64(Ops *t) {
  t[ops_key(2,2,1,2,2,0)] = make_ops<Cfg<2,2,true,2,2>, double>();
  t[ops_key(2,2,1,4,2,0)] = make_ops<Cfg<2,2,true,4,2>, double>();
}
}  // namespace idfx
