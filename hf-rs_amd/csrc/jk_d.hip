// jk_d.hip -- J/K kernels of the class (dd|dp) (36 output slices), a
// translation unit of its own so that the sliced classes compile in parallel.
#include "jk_kernel.h"

namespace hfg {

bool launch_jk_d(int cls, int v, const JKArgs& a, int max_blocks, hipStream_t s) {
  if (cls != 19) return false;
  launch_jk_v<2, 2, 2, 1>(v, a, max_blocks, s);
  return true;
}

}  // namespace hfg
