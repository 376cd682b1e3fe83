// jk_c.hip -- J/K kernels of the classes (dd|pp) and (dd|ds), and the class
// dispatch over the translation units jk_a .. jk_e.
#include "jk_kernel.h"

namespace hfg {

bool launch_jk_a(int cls, int v, const JKArgs& a, int max_blocks, hipStream_t s);
bool launch_jk_b(int cls, int v, const JKArgs& a, int max_blocks, hipStream_t s);
bool launch_jk_d(int cls, int v, const JKArgs& a, int max_blocks, hipStream_t s);
bool launch_jk_e(int cls, int v, const JKArgs& a, int max_blocks, hipStream_t s);

static bool launch_jk_c(int cls, int v, const JKArgs& a, int max_blocks, hipStream_t s) {
  switch (cls) {
    case 17: launch_jk_v<2, 2, 1, 1>(v, a, max_blocks, s); return true;
    case 18: launch_jk_v<2, 2, 2, 0>(v, a, max_blocks, s); return true;
    default: return false;
  }
}

// The case labels above follow quartet_class() order; checked here.
static_assert(class_index<0, 0, 0, 0>() == 0 && class_index<1, 1, 0, 0>() == 3 && class_index<2, 0, 2, 0>() == 9 &&
                  class_index<2, 1, 0, 0>() == 10 && class_index<2, 2, 1, 0>() == 16 &&
                  class_index<2, 2, 1, 1>() == 17 && class_index<2, 2, 2, 2>() == 20,
              "class order");

void launch_jk_class(int cls, int v, const JKArgs& a, int max_blocks, hipStream_t s) {
  if (!launch_jk_a(cls, v, a, max_blocks, s) && !launch_jk_b(cls, v, a, max_blocks, s) &&
      !launch_jk_c(cls, v, a, max_blocks, s) && !launch_jk_d(cls, v, a, max_blocks, s) &&
      !launch_jk_e(cls, v, a, max_blocks, s))
    throw HfgError{HFG_UNSUPPORTED, "no J/K kernel for class " + std::to_string(cls)};
}

}  // namespace hfg
