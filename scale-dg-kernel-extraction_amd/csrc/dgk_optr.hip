// The rest of the OPT1 operator family as batched device operators
// (mod_dg_optr_kernel_opt1.F90.erb): the directional kernels dirX/dirY/dirZ
// (:51-129), gradlike_dirXYZ (:166-219) and divlike_dirXYZ_var5 (:347-465),
// each over a trailing element dimension, one element per workgroup with the
// element's block and the matrix staged in LDS.
#include "dgk_internal.hpp"
#include "dgk_kernels.hpp"

using namespace dgk;

#define HIPOK(expr)                                                            \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess)                                                      \
      return fail(DGK_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

namespace {

// out(i,j,k) = sum_l Mat(i,l) in(l,j,k) (x), in(i,l,k) Mat(l,j) (y) or
// in(i,j,l) Mat(l,k) (z), left to right (erb:63-66, 92-95, 121-124).
template <int NQ>
__device__ __forceinline__ double dir_apply(int dir, const double *sM, const double *sx, int i,
                                            int j, int k) {
  constexpr int NFP = NQ * NQ;
  if (dir == 0)
    return contract<NQ, kSequential, false>([&](int l) { return sM[i + l * NQ]; },
                                            [&](int l) { return sx[l + j * NQ + k * NFP]; });
  if (dir == 1)
    return contract<NQ, kSequential, false>([&](int l) { return sx[i + l * NQ + k * NFP]; },
                                            [&](int l) { return sM[l + j * NQ]; });
  return contract<NQ, kSequential, false>([&](int l) { return sx[i + j * NQ + l * NFP]; },
                                          [&](int l) { return sM[l + k * NQ]; });
}

template <int NQ>
__global__ void __launch_bounds__(256)
    dir_batched_kernel(int dir, const double *Mat, const double *in, double *out) {
  constexpr int NP = NQ * NQ * NQ, NFP = NQ * NQ;
  extern __shared__ double smem[];
  double *sx = smem, *sM = smem + NP;
  const long long base = (long long)blockIdx.x * NP;
  for (int t = threadIdx.x; t < NP; t += blockDim.x) sx[t] = in[base + t];
  for (int t = threadIdx.x; t < NFP; t += blockDim.x) sM[t] = Mat[t];
  __syncthreads();
  for (int n = threadIdx.x; n < NP; n += blockDim.x) {
    const int i = n % NQ, j = (n / NQ) % NQ, k = n / NFP;
    out[base + n] = dir_apply<NQ>(dir, sM, sx, i, j, k);
  }
}

// gradlike_dirXYZ as its interface states: the x, y and z derivatives of one
// field (Mat along x, Mat_tr along y and z).
template <int NQ>
__global__ void __launch_bounds__(256)
    gradlike_batched_kernel(const double *Mat, const double *Mat_tr, const double *in,
                            double *ox, double *oy, double *oz) {
  constexpr int NP = NQ * NQ * NQ, NFP = NQ * NQ;
  extern __shared__ double smem[];
  double *sx = smem, *sM = smem + NP, *sT = smem + NP + NFP;
  const long long base = (long long)blockIdx.x * NP;
  for (int t = threadIdx.x; t < NP; t += blockDim.x) sx[t] = in[base + t];
  for (int t = threadIdx.x; t < NFP; t += blockDim.x) {
    sM[t] = Mat[t];
    sT[t] = Mat_tr[t];
  }
  __syncthreads();
  for (int n = threadIdx.x; n < NP; n += blockDim.x) {
    const int i = n % NQ, j = (n / NQ) % NQ, k = n / NFP;
    ox[base + n] = dir_apply<NQ>(0, sM, sx, i, j, k);
    oy[base + n] = dir_apply<NQ>(1, sT, sx, i, j, k);
    oz[base + n] = dir_apply<NQ>(2, sT, sx, i, j, k);
  }
}

int check_common(int Nq, int nelem) {
  if (Nq < 2 || Nq > 16) return fail(DGK_ERR_ORDER, "PolyOrder %d outside 1..15", Nq - 1);
  if (nelem < 0) return fail(DGK_ERR_ARG, "negative nelem");
  return DGK_OK;
}

}  // namespace

extern "C" {

int dgk_matvec_dir_batched(int Nq, int nelem, int dir, const double *Mat, const double *in,
                           double *out, void *stream) {
  clear_error();
  int rc = check_common(Nq, nelem);
  if (rc) return rc;
  if (dir < 0 || dir > 2) return fail(DGK_ERR_ARG, "direction %d is not 0 (x), 1 (y) or 2 (z)", dir);
  if (nelem == 0) return DGK_OK;
  if (!Mat || !in || !out) return fail(DGK_ERR_ARG, "null pointer");
  const size_t lds = sizeof(double) * ((size_t)Nq * Nq * Nq + (size_t)Nq * Nq);
  using Fn = void (*)(int, const double *, const double *, double *);
  Fn fn = nullptr;
#define DGK_DIR_CASE(NQ) \
  case NQ: fn = dir_batched_kernel<NQ>; break;
  switch (Nq) {
    DGK_DIR_CASE(2) DGK_DIR_CASE(3) DGK_DIR_CASE(4) DGK_DIR_CASE(5) DGK_DIR_CASE(6)
    DGK_DIR_CASE(7) DGK_DIR_CASE(8) DGK_DIR_CASE(9) DGK_DIR_CASE(10) DGK_DIR_CASE(11)
    DGK_DIR_CASE(12) DGK_DIR_CASE(13) DGK_DIR_CASE(14) DGK_DIR_CASE(15) DGK_DIR_CASE(16)
  }
#undef DGK_DIR_CASE
  HIPOK(hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lds));
  hipLaunchKernelGGL(fn, dim3(nelem), dim3(256), lds, (hipStream_t)stream, dir, Mat, in, out);
  HIPOK(hipGetLastError());
  return DGK_OK;
}

int dgk_gradlike_dirXYZ_batched(int Nq, int nelem, const double *Mat, const double *Mat_tr,
                                const double *in, double *out_x, double *out_y, double *out_z,
                                void *stream) {
  clear_error();
  int rc = check_common(Nq, nelem);
  if (rc) return rc;
  if (nelem == 0) return DGK_OK;
  if (!Mat || !Mat_tr || !in || !out_x || !out_y || !out_z)
    return fail(DGK_ERR_ARG, "null pointer");
  const size_t lds = sizeof(double) * ((size_t)Nq * Nq * Nq + 2 * (size_t)Nq * Nq);
  using Fn = void (*)(const double *, const double *, const double *, double *, double *,
                      double *);
  Fn fn = nullptr;
#define DGK_GRAD_CASE(NQ) \
  case NQ: fn = gradlike_batched_kernel<NQ>; break;
  switch (Nq) {
    DGK_GRAD_CASE(2) DGK_GRAD_CASE(3) DGK_GRAD_CASE(4) DGK_GRAD_CASE(5) DGK_GRAD_CASE(6)
    DGK_GRAD_CASE(7) DGK_GRAD_CASE(8) DGK_GRAD_CASE(9) DGK_GRAD_CASE(10) DGK_GRAD_CASE(11)
    DGK_GRAD_CASE(12) DGK_GRAD_CASE(13) DGK_GRAD_CASE(14) DGK_GRAD_CASE(15) DGK_GRAD_CASE(16)
  }
#undef DGK_GRAD_CASE
  HIPOK(hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lds));
  hipLaunchKernelGGL(fn, dim3(nelem), dim3(256), lds, (hipStream_t)stream, Mat, Mat_tr, in,
                     out_x, out_y, out_z);
  HIPOK(hipGetLastError());
  return DGK_OK;
}

int dgk_divlike_dirXYZ_var5_batched(int Nq, int nelem, const double *Mat, const double *Mat_tr,
                                    const double *in_x, const double *in_y, const double *in_z,
                                    double *out_x, double *out_y, double *out_z, void *stream) {
  clear_error();
  int rc = check_common(Nq, nelem);
  if (rc) return rc;
  if (nelem > 0x7fffffff / 5) return fail(DGK_ERR_ARG, "nelem %d too large", nelem);
  // five consecutive (Np) fields per element = 5*nelem divlike blocks
  return dgk_divlike_dirXYZ_batched(Nq, 5 * nelem, DGK_TYPE_OPT1, Mat, Mat_tr, in_x, in_y, in_z,
                                    out_x, out_y, out_z, stream);
}

}  // extern "C"
