// Kernel instantiations for one group of configurations (generated layout, one
// translation unit per group so the groups compile in parallel): 3-D isothermal MHD with
// HLLD and PLM, FP64 and FP32 stencil arithmetic.
#include "launch.hpp"

namespace idfx {
void register_iso3(Ops *t) {
  t[ops_key(3,3,1,2,2,0,1)] = make_ops<Cfg<3,3,true,2,2,true>, double>();
  t[ops_key(3,3,1,2,2,1,1)] = make_ops<Cfg<3,3,true,2,2,true>, float>();
}
}  // namespace idfx
