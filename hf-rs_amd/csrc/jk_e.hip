// jk_e.hip -- J/K kernels of the class (dd|dd) (36 output slices), a
// translation unit of its own so that the sliced classes compile in parallel.
#include "jk_kernel.h"

namespace hfg {

bool launch_jk_e(int cls, int v, const JKArgs& a, int max_blocks, hipStream_t s) {
  if (cls != 20) return false;
  launch_jk_v<2, 2, 2, 2>(v, a, max_blocks, s);
  return true;
}

}  // namespace hfg
