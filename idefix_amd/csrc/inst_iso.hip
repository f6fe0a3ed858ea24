// Kernel instantiations for one group of configurations (generated layout, one
// translation unit per group so the groups compile in parallel).
#include "launch.hpp"

namespace idfx {
void register_iso(Ops *t) {
  t[ops_key(1,1,0,2,1,0,1)] = make_ops<Cfg<1,1,false,2,1,true>, double>();
  t[ops_key(1,1,0,2,3,0,1)] = make_ops<Cfg<1,1,false,2,3,true>, double>();
  t[ops_key(1,1,0,2,4,0,1)] = make_ops<Cfg<1,1,false,2,4,true>, double>();
  t[ops_key(1,1,0,2,5,0,1)] = make_ops<Cfg<1,1,false,2,5,true>, double>();
  t[ops_key(1,1,0,3,1,0,1)] = make_ops<Cfg<1,1,false,3,1,true>, double>();
  t[ops_key(1,1,0,4,1,0,1)] = make_ops<Cfg<1,1,false,4,1,true>, double>();
  t[ops_key(1,1,0,4,3,0,1)] = make_ops<Cfg<1,1,false,4,3,true>, double>();
  t[ops_key(2,2,0,2,1,0,1)] = make_ops<Cfg<2,2,false,2,1,true>, double>();
  t[ops_key(1,2,1,2,2,0,1)] = make_ops<Cfg<1,2,true,2,2,true>, double>();
  t[ops_key(1,2,1,2,4,0,1)] = make_ops<Cfg<1,2,true,2,4,true>, double>();
  t[ops_key(1,2,1,2,5,0,1)] = make_ops<Cfg<1,2,true,2,5,true>, double>();
  t[ops_key(1,2,1,3,2,0,1)] = make_ops<Cfg<1,2,true,3,2,true>, double>();
  t[ops_key(1,2,1,4,2,0,1)] = make_ops<Cfg<1,2,true,4,2,true>, double>();
  t[ops_key(2,2,1,2,2,0,1)] = make_ops<Cfg<2,2,true,2,2,true>, double>();
  t[ops_key(1,2,1,2,3,0,1)] = make_ops<Cfg<1,2,true,2,3,true>, double>();
  t[ops_key(1,2,1,3,3,0,1)] = make_ops<Cfg<1,2,true,3,3,true>, double>();
  t[ops_key(1,2,1,4,3,0,1)] = make_ops<Cfg<1,2,true,4,3,true>, double>();
  t[ops_key(2,2,1,2,3,0,1)] = make_ops<Cfg<2,2,true,2,3,true>, double>();
}
}  // namespace idfx
