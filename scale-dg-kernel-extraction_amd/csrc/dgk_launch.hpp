// Launch descriptors shared by the solver's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include "dgk_kernels.hpp"

namespace dgk {

// Device copies of the 1D operator in row-major order (D(k,l) at
// [k*Nq + l], read a row at a time by the pencil kernel) and of the vectors
// [l_-, l_+] of the factored lift, followed by the pairs (l_-(n), l_+(n)) for
// n < Nq (4 Nq doubles; the one-wavefront stage reads the pairs).
struct HostOps {
  const double *Drow;
  const double *lfac;
  const double *LiftP;   // pair-packed Lift for the pencil kernel
  const double *Drow32;  // mixed precision: split rows and line-fit tables (split_tables)
  const double *mfit;    // Nq = 16, matrix-core mixed form: line-fit tables (mfma_fit_tables)
};

struct StageKernel;
using LaunchFn = void (*)(const StageKernel &, const StageParams &, const HostOps &,
                          const double *Lift, hipStream_t);

struct StageKernel {
  const void *kernel = nullptr;  // for attributes and occupancy
  LaunchFn launch = nullptr;
  int nt = 0;
  size_t lds = 0;
  int epb = 1;       // elements per workgroup (node kernel, group pencil) or per
                     // wavefront (one-wavefront pencil)
  bool pencil = false;
  long long grid = 0;
};

// The multi-wavefront pencil stage (pencil_group_kernel, lean LDS layout)
// for Nq in 2..16 with epb elements per workgroup; an empty descriptor if
// epb does not fit the kernel's thread limit.
StageKernel pick_pencil_group(int Nq, int order, bool fma, bool fact, bool cv, int epb);
// the same for Nq 2-8, 9-12 and 13-16 (separate translation units)
StageKernel pick_pencil_group_low(int Nq, int order, bool fma, bool fact, bool cv, int epb);
StageKernel pick_pencil_group_mid(int Nq, int order, bool fma, bool fact, bool cv, int epb);
StageKernel pick_pencil_group_high(int Nq, int order, bool fma, bool fact, bool cv, int epb);
// The matrix-core stage (dgk_mfma.hip) for Nq = 16 with uniform velocity:
// FP64 MFMA contractions, or with `mixed` FP32 MFMA contractions of each
// line's remainder after an FP64 degree-7 fit; `wide`: eight wavefronts per
// element instead of four.  An empty descriptor for other Nq.
StageKernel pick_mfma(int Nq, bool mixed, bool wide);

}  // namespace dgk
