// Kernel instantiations for one group of configurations (generated layout, one
// translation unit per group so the groups compile in parallel).
#include "launch.hpp"

namespace idfx {
void register_mhd1(Ops *t) {
  t[ops_key(1,2,1,2,2,0)] = make_ops<Cfg<1,2,true,2,2>, double>();
  t[ops_key(1,2,1,2,3,0)] = make_ops<Cfg<1,2,true,2,3>, double>();
  t[ops_key(1,2,1,2,4,0)] = make_ops<Cfg<1,2,true,2,4>, double>();
  t[ops_key(1,2,1,2,5,0)] = make_ops<Cfg<1,2,true,2,5>, double>();
  t[ops_key(1,2,1,3,2,0)] = make_ops<Cfg<1,2,true,3,2>, double>();
  t[ops_key(1,2,1,3,3,0)] = make_ops<Cfg<1,2,true,3,3>, double>();
  t[ops_key(1,2,1,3,4,0)] = make_ops<Cfg<1,2,true,3,4>, double>();
  t[ops_key(1,2,1,3,5,0)] = make_ops<Cfg<1,2,true,3,5>, double>();
  t[ops_key(1,2,1,4,2,0)] = make_ops<Cfg<1,2,true,4,2>, double>();
  t[ops_key(1,2,1,4,3,0)] = make_ops<Cfg<1,2,true,4,3>, double>();
}
}  // namespace idfx
