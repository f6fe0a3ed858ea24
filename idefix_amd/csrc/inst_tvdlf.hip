// Kernel instantiations for one group of configurations (generated layout, one
// translation unit per group so the groups compile in parallel).
#include "launch.hpp"

namespace idfx {
void register_tvdlf(Ops *t) {
  t[ops_key(1,1,0,2,5,0)] = make_ops<Cfg<1,1,false,2,5>, double>();
  t[ops_key(2,2,0,2,5,0)] = make_ops<Cfg<2,2,false,2,5>, double>();
  t[ops_key(3,3,0,2,5,0)] = make_ops<Cfg<3,3,false,2,5>, double>();
  t[ops_key(2,2,1,2,5,0)] = make_ops<Cfg<2,2,true,2,5>, double>();
  t[ops_key(3,3,1,2,5,0)] = make_ops<Cfg<3,3,true,2,5>, double>();
}
}  // namespace idfx
