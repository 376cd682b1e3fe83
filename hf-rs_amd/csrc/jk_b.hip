// jk_b.hip -- J/K kernels of the classes (dp|ss) .. (dd|ps).
#include "jk_kernel.h"

namespace hfg {

bool launch_jk_b(int cls, int v, const JKArgs& a, int max_blocks, hipStream_t s) {
  switch (cls) {
    case 10: launch_jk_v<2, 1, 0, 0>(v, a, max_blocks, s); return true;
    case 11: launch_jk_v<2, 1, 1, 0>(v, a, max_blocks, s); return true;
    case 12: launch_jk_v<2, 1, 1, 1>(v, a, max_blocks, s); return true;
    case 13: launch_jk_v<2, 1, 2, 0>(v, a, max_blocks, s); return true;
    case 14: launch_jk_v<2, 1, 2, 1>(v, a, max_blocks, s); return true;
    case 15: launch_jk_v<2, 2, 0, 0>(v, a, max_blocks, s); return true;
    case 16: launch_jk_v<2, 2, 1, 0>(v, a, max_blocks, s); return true;
    default: return false;
  }
}

}  // namespace hfg
