// Kernel instantiations for one group of configurations (generated layout, one
// translation unit per group so the groups compile in parallel).
#include "launch.hpp"

namespace idfx {
void register_mlp(Ops *t) {
  t[ops_key(2,3,1,8,2,0)] = make_ops<Cfg<2,3,true,8,2>, double>();
  t[ops_key(2,3,1,8,2,1)] = make_ops<Cfg<2,3,true,8,2>, float>();
  t[ops_key(3,3,1,8,2,0)] = make_ops<Cfg<3,3,true,8,2>, double>();
  t[ops_key(3,3,1,8,2,1)] = make_ops<Cfg<3,3,true,8,2>, float>();
}
}  // namespace idfx
