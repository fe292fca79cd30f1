// Instantiation of the multi-wavefront pencil stage (pencil_group_kernel) for
// one Nq: shared by the translation units that build the orders in parallel.
#pragma once

#include "dgk_launch.hpp"
#include "dgk_pencil.hpp"

namespace dgk {
namespace group_impl {

// Workgroup limit per Nq: Nq 9-12 and 13-16 need about 150-230 VGPRs, so their
// workgroups stay at five and four wavefronts (at most two per SIMD, which
// allows 256 VGPRs); lower orders allow nine.
template <int NQ>
constexpr int group_maxth() { return NQ >= 13 ? 256 : (NQ >= 9 ? 320 : 576); }

template <int NQ, int ORDER, bool FMA, bool FACT, bool CV>
void launch_group(const StageKernel &k, const StageParams &P, const HostOps &H, const double *,
                  hipStream_t st) {
  // the FP32 split form reads the split copies of D and of the lift vectors
  const double *dr = ORDER == kMixed ? H.Drow32 : H.Drow;
  const double *lf = H.lfac;
  hipLaunchKernelGGL((pencil_group_kernel<NQ, group_maxth<NQ>(), ORDER, FMA, FACT, true, CV>),
                     dim3((unsigned)k.grid), dim3(k.nt), k.lds, st, P, dr, H.LiftP, lf, k.epb);
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
  if (order == kMixed) {  // FP32 split contractions: plain products, factored lift only
    if (fma || !fact) return StageKernel{};
    return cv ? make_group<NQ, kMixed, false, true, true>(epb)
              : make_group<NQ, kMixed, false, true, false>(epb);
  }
  if (order == kGeneral)
    return fma ? pick_fact<NQ, kGeneral, true>(fact, cv, epb)
               : pick_fact<NQ, kGeneral, false>(fact, cv, epb);
  return fma ? pick_fact<NQ, kOpt1, true>(fact, cv, epb)
             : pick_fact<NQ, kOpt1, false>(fact, cv, epb);
}


}  // namespace group_impl
}  // namespace dgk
