// Kernel instantiations for one group of configurations (generated layout, one
// translation unit per group so the groups compile in parallel).
#include "launch.hpp"

namespace idfx {
void register_limo3_mhd_b(Ops *t) {
  t[ops_key(2,2,1,3,5,0)] = make_ops<Cfg<2,2,true,3,5>, double>();
  t[ops_key(2,2,1,4,4,0)] = make_ops<Cfg<2,2,true,4,4>, double>();
  t[ops_key(2,2,1,4,5,0)] = make_ops<Cfg<2,2,true,4,5>, double>();
}
}  // namespace idfx
