// Kernel instantiations for one group of configurations (generated layout, one
// translation unit per group so the groups compile in parallel).
#include "launch.hpp"

namespace idfx {
void register_limo3_mhd(Ops *t) {
  t[ops_key(2,2,1,3,2,0)] = make_ops<Cfg<2,2,true,3,2>, double>();
  t[ops_key(2,2,1,3,4,0)] = make_ops<Cfg<2,2,true,3,4>, double>();
  t[ops_key(3,3,1,3,2,0)] = make_ops<Cfg<3,3,true,3,2>, double>();
}
}  // namespace idfx
