// Kernel instantiations for one group of configurations (generated layout, one
// translation unit per group so the groups compile in parallel).
#include "launch.hpp"

namespace idfx {
void register_mhd2c3(Ops *t) {
  t[ops_key(2,3,1,2,2,0)] = make_ops<Cfg<2,3,true,2,2>, double>();
  t[ops_key(2,3,1,4,2,0)] = make_ops<Cfg<2,3,true,4,2>, double>();
  t[ops_key(2,3,1,2,2,1)] = make_ops<Cfg<2,3,true,2,2>, float>();
  t[ops_key(2,3,1,4,2,1)] = make_ops<Cfg<2,3,true,4,2>, float>();
}
}  // namespace idfx
