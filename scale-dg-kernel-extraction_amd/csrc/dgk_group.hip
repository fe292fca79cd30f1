// The multi-wavefront pencil stage (pencil_group_kernel) for every pencil
// order, built in a translation unit of its own.
#include "dgk_launch.hpp"
#include "dgk_pencil.hpp"

namespace dgk {
namespace {

// Workgroup limit per Nq: Nq 12 and 16 need about 150-230 VGPRs, so their
// workgroups stay at five and four wavefronts (at most two per SIMD, which
// allows 256 VGPRs); the other orders allow nine.
template <int NQ>
constexpr int group_maxth() { return NQ == 16 ? 256 : (NQ == 12 ? 320 : 576); }

template <int NQ, int ORDER, bool FMA, bool FACT, bool CV>
void launch_group(const StageKernel &k, const StageParams &P, const HostOps &H, const double *,
                  hipStream_t st) {
  hipLaunchKernelGGL((pencil_group_kernel<NQ, group_maxth<NQ>(), ORDER, FMA, FACT, true, CV>),
                     dim3((unsigned)k.grid), dim3(k.nt), k.lds, st, P, H.Drow, H.LiftP, H.lfac,
                     k.epb);
}

template <int NQ, int ORDER, bool FMA, bool FACT, bool CV>
StageKernel make_group(int epb) {
  StageKernel k;
  if (epb < 1 || pencil_group_threads<NQ>(epb) > group_maxth<NQ>()) return k;
  k.kernel = (const void *)pencil_group_kernel<NQ, group_maxth<NQ>(), ORDER, FMA, FACT, true, CV>;
  k.launch = launch_group<NQ, ORDER, FMA, FACT, CV>;
  k.epb = epb;
  k.nt = pencil_group_threads<NQ>(epb);
  k.lds = pencil_group_lds<NQ>(epb, CV);
  k.pencil = true;
  return k;
}

template <int NQ, int ORDER, bool FMA>
StageKernel pick_fact(bool fact, bool cv, int epb) {
  if (fact)
    return cv ? make_group<NQ, ORDER, FMA, true, true>(epb)
              : make_group<NQ, ORDER, FMA, true, false>(epb);
  return cv ? make_group<NQ, ORDER, FMA, false, true>(epb)
            : make_group<NQ, ORDER, FMA, false, false>(epb);
}

template <int NQ>
StageKernel pick_nq(int order, bool fma, bool fact, bool cv, int epb) {
  if (order == kGeneral)
    return fma ? pick_fact<NQ, kGeneral, true>(fact, cv, epb)
               : pick_fact<NQ, kGeneral, false>(fact, cv, epb);
  return fma ? pick_fact<NQ, kOpt1, true>(fact, cv, epb)
             : pick_fact<NQ, kOpt1, false>(fact, cv, epb);
}

}  // namespace

StageKernel pick_pencil_group(int Nq, int order, bool fma, bool fact, bool cv, int epb) {
  switch (Nq) {
  case 2: return pick_nq<2>(order, fma, fact, cv, epb);
  case 4: return pick_nq<4>(order, fma, fact, cv, epb);
  case 6: return pick_nq<6>(order, fma, fact, cv, epb);
  case 8: return pick_nq<8>(order, fma, fact, cv, epb);
  case 12: return pick_nq<12>(order, fma, fact, cv, epb);
  case 16: return pick_nq<16>(order, fma, fact, cv, epb);
  default: return StageKernel{};
  }
}

}  // namespace dgk
