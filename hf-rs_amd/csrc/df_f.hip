// df_f.hip -- block kernels of the density-fitting 3-index integrals
// (P|mu nu) (three_center_integrals, src/df.rs:60-78) whose orbital pair
// holds an f shell with a p, d or f partner: (fp|X 1), (fd|X 1), (ff|X 1) for
// an auxiliary shell X up to g with the unit function.  Generated HGP code
// (gen/eri_df.inc), one thread per block as the engine's block kernels, but
// each block written straight to its output slot: at up to 1,500 doubles a
// block a private copy would only add scratch.
#include <string>

#include "engine.h"
#include "eri_device.h"

namespace hfg {
#include "gen/eri_df.inc"

namespace {

template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(64) void k_dfblock(BlockArgs a) {
  __shared__ double boys[BoysSlice<false>::size];
  load_boys_slice<LA + LB + LC + LD>(a.boys, boys);
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.ntask) return;
  const BlockTask bt = a.tasks[t];
  const PairDev& pb = a.pairs[bt.bra];
  const PairDev& pk = a.pairs[bt.ket];
  eri_block<LA, LB, LC, LD>(a.pp, pb.p0, pb.p1, a.pp, pk.p0, pk.p1 - pk.p0, 1, pb.AB, pk.AB, boys, a.out + bt.off);
}

}  // namespace

void launch_df_f_block(int l1, int l2, int l3, int l4, const BlockArgs& a, hipStream_t s) {
#define X(p, q, r, u)                                                                            \
  if (p == l1 && q == l2 && r == l3 && u == l4) {                                                  \
    hipLaunchKernelGGL((k_dfblock<p, q, r, u>), dim3((a.ntask + 63) / 64), dim3(64), 0, s, a);      \
    HIP_CHECK(hipGetLastError());                                                                  \
    return;                                                                                        \
  }
  HFG_FOR_EACH_DF_F_CLASS(X)
#undef X
  throw HfgError{HFG_UNSUPPORTED, "no 3-index block kernel for class (" + std::to_string(l1) + std::to_string(l2) +
                                      "|" + std::to_string(l3) + std::to_string(l4) + ")"};
}

}  // namespace hfg
