// The multi-wavefront pencil stage for Nq 9, 10, 11, 12.
#include "dgk_group_impl.hpp"

namespace dgk {

StageKernel pick_pencil_group_mid(int Nq, int order, bool fma, bool fact, bool cv, int epb) {
  switch (Nq) {
  case 9: return group_impl::pick_nq<9>(order, fma, fact, cv, epb);
  case 10: return group_impl::pick_nq<10>(order, fma, fact, cv, epb);
  case 11: return group_impl::pick_nq<11>(order, fma, fact, cv, epb);
  case 12: return group_impl::pick_nq<12>(order, fma, fact, cv, epb);
  default: return StageKernel{};
  }
}

}  // namespace dgk
