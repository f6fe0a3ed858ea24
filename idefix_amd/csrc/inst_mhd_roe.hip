// Kernel instantiations for one group of configurations (generated layout, one
// translation unit per group so the groups compile in parallel).
#include "launch.hpp"

namespace idfx {
void register_mhd_roe(Ops *t) {
  t[ops_key(2,2,1,2,3,0)] = make_ops<Cfg<2,2,true,2,3>, double>();
  t[ops_key(2,2,1,4,3,0)] = make_ops<Cfg<2,2,true,4,3>, double>();
  t[ops_key(3,3,1,2,3,0)] = make_ops<Cfg<3,3,true,2,3>, double>();
  t[ops_key(3,3,1,4,3,0)] = make_ops<Cfg<3,3,true,4,3>, double>();
}
}  // namespace idfx
