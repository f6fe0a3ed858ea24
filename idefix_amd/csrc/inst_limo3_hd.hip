// Kernel instantiations for one group of configurations (generated layout, one
// translation unit per group so the groups compile in parallel).
#include "launch.hpp"

namespace idfx {
void register_limo3_hd(Ops *t) {
  t[ops_key(1,1,0,3,1,0)] = make_ops<Cfg<1,1,false,3,1>, double>();
  t[ops_key(1,1,0,3,3,0)] = make_ops<Cfg<1,1,false,3,3>, double>();
  t[ops_key(1,1,0,3,4,0)] = make_ops<Cfg<1,1,false,3,4>, double>();
  t[ops_key(1,1,0,3,5,0)] = make_ops<Cfg<1,1,false,3,5>, double>();
  t[ops_key(1,1,0,4,4,0)] = make_ops<Cfg<1,1,false,4,4>, double>();
  t[ops_key(1,1,0,4,5,0)] = make_ops<Cfg<1,1,false,4,5>, double>();
  t[ops_key(2,2,0,3,1,0)] = make_ops<Cfg<2,2,false,3,1>, double>();
}
}  // namespace idfx
