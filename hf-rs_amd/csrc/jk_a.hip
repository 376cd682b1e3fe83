// jk_a.hip -- J/K kernels of the classes (ss|ss) .. (ds|ds); jk_b.hip and
// jk_c.hip hold the rest, so that the three compile in parallel.
#include "jk_kernel.h"

namespace hfg {

bool launch_jk_a(int cls, int v, const JKArgs& a, int max_blocks, hipStream_t s) {
  switch (cls) {
    case 0: launch_jk_v<0, 0, 0, 0>(v, a, max_blocks, s); return true;
    case 1: launch_jk_v<1, 0, 0, 0>(v, a, max_blocks, s); return true;
    case 2: launch_jk_v<1, 0, 1, 0>(v, a, max_blocks, s); return true;
    case 3: launch_jk_v<1, 1, 0, 0>(v, a, max_blocks, s); return true;
    case 4: launch_jk_v<1, 1, 1, 0>(v, a, max_blocks, s); return true;
    case 5: launch_jk_v<1, 1, 1, 1>(v, a, max_blocks, s); return true;
    case 6: launch_jk_v<2, 0, 0, 0>(v, a, max_blocks, s); return true;
    case 7: launch_jk_v<2, 0, 1, 0>(v, a, max_blocks, s); return true;
    case 8: launch_jk_v<2, 0, 1, 1>(v, a, max_blocks, s); return true;
    case 9: launch_jk_v<2, 0, 2, 0>(v, a, max_blocks, s); return true;
    default: return false;
  }
}

}  // namespace hfg
