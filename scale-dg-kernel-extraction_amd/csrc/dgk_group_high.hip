// The multi-wavefront pencil stage for Nq 13, 14, 15, 16.
#include "dgk_group_impl.hpp"

namespace dgk {

StageKernel pick_pencil_group_high(int Nq, int order, bool fma, bool fact, bool cv, int epb) {
  switch (Nq) {
  case 13: return group_impl::pick_nq<13>(order, fma, fact, cv, epb);
  case 14: return group_impl::pick_nq<14>(order, fma, fact, cv, epb);
  case 15: return group_impl::pick_nq<15>(order, fma, fact, cv, epb);
  case 16: return group_impl::pick_nq<16>(order, fma, fact, cv, epb);
  default: return StageKernel{};
  }
}

}  // namespace dgk
