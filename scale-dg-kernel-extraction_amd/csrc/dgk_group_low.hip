// The multi-wavefront pencil stage (pencil_group_kernel) for Nq 2-8, and the
// dispatch over the three translation units that build it.
#include "dgk_group_impl.hpp"

namespace dgk {

StageKernel pick_pencil_group_low(int Nq, int order, bool fma, bool fact, bool cv, int epb) {
  switch (Nq) {
  case 2: return group_impl::pick_nq<2>(order, fma, fact, cv, epb);
  case 3: return group_impl::pick_nq<3>(order, fma, fact, cv, epb);
  case 4: return group_impl::pick_nq<4>(order, fma, fact, cv, epb);
  case 5: return group_impl::pick_nq<5>(order, fma, fact, cv, epb);
  case 6: return group_impl::pick_nq<6>(order, fma, fact, cv, epb);
  case 7: return group_impl::pick_nq<7>(order, fma, fact, cv, epb);
  case 8: return group_impl::pick_nq<8>(order, fma, fact, cv, epb);
  default: return StageKernel{};
  }
}

StageKernel pick_pencil_group(int Nq, int order, bool fma, bool fact, bool cv, int epb) {
  if (Nq <= 8) return pick_pencil_group_low(Nq, order, fma, fact, cv, epb);
  if (Nq <= 12) return pick_pencil_group_mid(Nq, order, fma, fact, cv, epb);
  return pick_pencil_group_high(Nq, order, fma, fact, cv, epb);
}

}  // namespace dgk
