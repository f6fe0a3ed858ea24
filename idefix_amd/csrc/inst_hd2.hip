// Kernel instantiations for one group of configurations (generated layout, one
// translation unit per group so the groups compile in parallel).
#include "launch.hpp"

namespace idfx {
void register_hd2(Ops *t) {
  t[ops_key(2,2,0,2,1,0)] = make_ops<Cfg<2,2,false,2,1>, double>();
  t[ops_key(2,2,0,2,3,0)] = make_ops<Cfg<2,2,false,2,3>, double>();
  t[ops_key(2,2,0,4,1,0)] = make_ops<Cfg<2,2,false,4,1>, double>();
  t[ops_key(2,2,0,4,3,0)] = make_ops<Cfg<2,2,false,4,3>, double>();
}
}  // namespace idfx
