// Matrix-core form of the fused RK stage for Nq = 16 (p = 15) with uniform
// velocity: the three directional contractions run as 16x16x4 MFMA products
// of the 1D operator with the element's nodal block, in FP64
// (v_mfma_f64_16x16x4_f64) or, for DGK_FLAG_MIXED, in FP32 with operator and
// operand split into hi + lo parts (three v_mfma_f32_16x16x4_f32 per product:
// D_hi q_hi + D_lo q_hi + D_hi q_lo, one FP32 accumulator).  Face fluxes,
// the factored lift, the residual sum and the RK update are FP64.
//
// One element per workgroup of NW = 4 or 8 wavefronts.  Wavefront w, lane
// L = 16 g + c holds in registers the nodes (i, j, k) = (c, BW w + b, k_r)
// for its BW = 16 / NW blocks b and the accumulator rows r = 0..3, with
// k_r = g + 4r for the FP64
// product and 4g + r for the FP32 one (the rows each MFMA form returns to a
// lane).  The block q(Np) is staged once in LDS; the x and z contractions
// read it as the MFMA operand of their k-by-i products and land directly in
// that register layout, both accumulated into the same registers (the
// scale factors Es_d u_d folded into the operator); the y contraction's
// k-by-j product goes back through LDS.  Three barriers per element.
//
// MFMA operand layout of the 16x16x4 forms (checked on the device): lane L
// supplies A[m = L % 16][kk = L / 16] and B[kk = L / 16][n = L % 16]; the
// result D[m][n] sits in lane n + 16 g as D[g + 4r][n] (FP64) or
// D[4g + r][n] (FP32), r = 0..3.
#include "dgk_launch.hpp"
#include "dgk_pencil.hpp"

namespace dgk {
namespace {

constexpr int kMq = 16, kMfp = kMq * kMq, kMp = kMq * kMq * kMq;
// LDS position of node (i, j, k): one pad slot per x-line and one per
// xy-plane, so that the 16 lanes of a row group hit distinct banks whether
// they walk i (stride 1), j (stride 17) or k (stride 273).
__device__ __forceinline__ int qpos(int i, int j, int k) { return i + 17 * j + 273 * k; }
constexpr int kQSlots = 15 + 17 * 15 + 273 * 15 + 1;  // 4366
constexpr size_t kMfmaLds = sizeof(double) * (kQSlots + 6 * kMfp);

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

template <bool F32>
struct MfmaAcc;
template <>
struct MfmaAcc<false> {
  d4 v;
  __device__ __forceinline__ void zero() { v = d4{0.0, 0.0, 0.0, 0.0}; }
  __device__ __forceinline__ double get(int r) const { return v[r]; }
};
template <>
struct MfmaAcc<true> {
  f4 v;
  __device__ __forceinline__ void zero() { v = f4{0.0f, 0.0f, 0.0f, 0.0f}; }
  __device__ __forceinline__ double get(int r) const { return (double)v[r]; }
};

// An LDS operand: the FP64 value, or its (hi, lo) FP32 split.
template <bool F32>
using Operand = typename std::conditional<F32, float2, double>::type;

// acc += A B for one 16x16x4 step; a and b are (hi, lo) pairs in FP32.
__device__ __forceinline__ void mma(MfmaAcc<false> &acc, double a, double b) {
  acc.v = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc.v, 0, 0, 0);
}
__device__ __forceinline__ void mma(MfmaAcc<true> &acc, float2 a, float2 b) {
  acc.v = __builtin_amdgcn_mfma_f32_16x16x4f32(a.x, b.x, acc.v, 0, 0, 0);
  acc.v = __builtin_amdgcn_mfma_f32_16x16x4f32(a.y, b.x, acc.v, 0, 0, 0);
  acc.v = __builtin_amdgcn_mfma_f32_16x16x4f32(a.x, b.y, acc.v, 0, 0, 0);
}

__device__ __forceinline__ double to_op(double x, std::false_type) { return x; }
__device__ __forceinline__ float2 to_op(double x, std::true_type) {
  const float hi = (float)x;
  return float2{hi, (float)(x - (double)hi)};
}

// The FP32 four-wavefront form fits three wavefronts per SIMD without
// spilling (measured 2.5 % faster than two); the FP64 form spills there
// (47 % slower) and keeps two.
template <bool F32, int NW>
__global__ void __launch_bounds__(64 * NW, F32 && NW == 4 ? 3 : 1)
    mfma16_kernel(const StageParams P, const double *__restrict__ Drow,
                  const double *__restrict__ lfac) {
  constexpr int BW = kMq / NW;  // blocks (j, or i for y) per wavefront
  using Op = Operand<F32>;
  using IsF32 = std::integral_constant<bool, F32>;
  extern __shared__ double smem[];
  Op *Q = reinterpret_cast<Op *>(smem);  // q, then the y contraction (as double)
  double *Y = smem;
  double *F = smem + kQSlots;            // six face-flux planes of 256
  const int tid = threadIdx.x, w = tid >> 6, L = tid & 63, c = L & 15, g = L >> 4;

  // element of this workgroup (XCD-aware order, as the pencil kernels)
  const int xcd = blockIdx.x % kNumXcd;
  const long long gi = blockIdx.x / kNumXcd;
  long long e;
  if (P.tnx) {
    e = tiled_lookup(P, xcd, gi);
  } else {
    const long long per = (P.Ne + kNumXcd - 1) / kNumXcd;
    e = gi < per ? (long long)xcd * per + gi : P.Ne;
    if (e >= P.Ne) e = -1;
  }
  if (e < 0) return;  // uniform over the workgroup
  auto krow = [&](int r) { return F32 ? 4 * g + r : g + 4 * r; };
  const long long eb = e * kMp;
  const int32_t *nb = P.nbr + e * 6;

  // ---- loads: own nodes, and the neighbour traces of the face nodes held --
  // q0 (stages 2 and 3) with q, so that its latency is hidden by the stage
  double q[BW][4], qz0[BW][4];
  const bool need_q0 = P.mode == kModeStageN;
#pragma unroll
  for (int b = 0; b < BW; ++b)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const long long idx = eb + c + kMq * (BW * w + b) + kMfp * krow(r);
      q[b][r] = P.qin[idx];
      qz0[b][r] = need_q0 ? ld_stream(P.q0 + idx) : 0.0;
    }
  const double cu = P.cvel[0], cv = P.cvel[1], cw = P.cvel[2];
  // x faces (i = 0: face 3, -x; i = 15: face 1, +x) at (j, k)
  if (c == 0 || c == kMq - 1) {
    const int f = c == 0 ? 3 : 1;
    const long long nbase = (long long)nb[f] * kMp + (c == 0 ? kMq - 1 : 0);
    const double vm = c == 0 ? -cu : cu;
#pragma unroll
    for (int b = 0; b < BW; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int j = BW * w + b, k = krow(r);
        const double qp = P.qin[nbase + kMq * j + kMfp * k];
        F[f * kMfp + j + kMq * k] = upwind_flux(q[b][r], qp, vm, vm, P.fs[f]);
      }
  }
  // y faces (j = 0: face 0, -y; j = 15: face 2, +y) at (i, k)
  if ((w == 0) || (w == NW - 1)) {
    const int b = w == 0 ? 0 : BW - 1, f = w == 0 ? 0 : 2;
    const long long nbase = (long long)nb[f] * kMp + c + kMq * (w == 0 ? kMq - 1 : 0);
    const double vm = w == 0 ? -cv : cv;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int k = krow(r);
      const double qp = P.qin[nbase + kMfp * k];
      F[f * kMfp + c + kMq * k] = upwind_flux(q[b][r], qp, vm, vm, P.fs[f]);
    }
  }
  // z faces (k = 0: face 4, -z; k = 15: face 5, +z) at (i, j): rows r = 0 of
  // g = 0 and r = 3 of g = 3 in both row layouts
  if (g == 0 || g == 3) {
    const int r = g == 0 ? 0 : 3, f = g == 0 ? 4 : 5;
    const long long nbase = (long long)nb[f] * kMp + c + kMfp * (g == 0 ? kMq - 1 : 0);
    const double vm = g == 0 ? -cw : cw;
#pragma unroll
    for (int b = 0; b < BW; ++b) {
      const int j = BW * w + b;
      const double qp = P.qin[nbase + kMq * j];
      F[f * kMfp + c + kMq * j] = upwind_flux(q[b][r], qp, vm, vm, P.fs[f]);
    }
  }
#pragma unroll
  for (int b = 0; b < BW; ++b)
#pragma unroll
    for (int r = 0; r < 4; ++r) Q[qpos(c, BW * w + b, krow(r))] = to_op(q[b][r], IsF32{});

  // operator operands: D(c, 4s + g), scaled by Es_d u_d of each direction
  Op ax[4], ay[4], az[4];
  {
    const double sx = P.es[0] * cu, sy = P.es[1] * cv, sz = P.es[2] * cw;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const double d = Drow[kMq * c + 4 * s + g];
      ax[s] = to_op(sx * d, IsF32{});
      ay[s] = to_op(sy * d, IsF32{});
      az[s] = to_op(sz * d, IsF32{});
    }
  }
  __syncthreads();

  // ---- contractions -------------------------------------------------------
  // x: out(i,j,k) = sum_l D(i,l) q(l,j,k): A[m=k][kk=l] = q(l,j,k),
  //    B[kk=l][n=i] = D(i,l); z: out = sum_l D(k,l) q(i,j,l): A = D(k,l),
  //    B[kk=l][n=i] = q(i,j,l); both k-by-i for j = 4w + b.
  // y: out(i,j,k) = sum_l D(j,l) q(i,l,k): A[m=k][kk=l] = q(i,l,k),
  //    B[kk=l][n=j] = D(j,l), k-by-j for i = 4w + b.
  MfmaAcc<F32> acc[BW], accy[BW];
#pragma unroll
  for (int b = 0; b < BW; ++b) {
    acc[b].zero();
    accy[b].zero();
    const int jb = BW * w + b;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int l = 4 * s + g;
      mma(acc[b], Q[qpos(l, jb, c)], ax[s]);
      mma(acc[b], az[s], Q[qpos(c, jb, l)]);
      mma(accy[b], Q[qpos(jb, l, c)], ay[s]);
    }
  }
  __syncthreads();  // every read of the staged q is done
  // the y contraction at (i = 4w + b, j = c, k_r), over the staged q
#pragma unroll
  for (int b = 0; b < BW; ++b)
#pragma unroll
    for (int r = 0; r < 4; ++r) Y[qpos(BW * w + b, c, krow(r))] = accy[b].get(r);
  __syncthreads();

  // ---- lift, residual, RK update, stores ------------------------------------
  // factored lift: l_-(normal index) for faces -y, -x, -z; l_+ for +x, +y,
  // +z; lfac = [l_-, l_+]
  const double lmi = lfac[c], lpi = lfac[kMq + c];
#pragma unroll
  for (int b = 0; b < BW; ++b) {
    const int j = BW * w + b;
    const double lmj = lfac[j], lpj = lfac[kMq + j];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int k = krow(r);
      double lift = lmj * F[0 * kMfp + c + kMq * k];
      lift = __builtin_fma(lpi, F[1 * kMfp + j + kMq * k], lift);
      lift = __builtin_fma(lpj, F[2 * kMfp + c + kMq * k], lift);
      lift = __builtin_fma(lmi, F[3 * kMfp + j + kMq * k], lift);
      lift = __builtin_fma(lfac[k], F[4 * kMfp + c + kMq * j], lift);
      lift = __builtin_fma(lfac[kMq + k], F[5 * kMfp + c + kMq * j], lift);
      const double dq = -((acc[b].get(r) + Y[qpos(c, j, k)]) + lift);
      const long long idx = eb + c + kMq * j + kMfp * k;
      if (P.mode == kModeTend) {
        st<true>(P.dqdt + idx, dq);
      } else {
        const double qq = q[b][r];
        const double qs = P.mode == kModeStage1 ? qq : qz0[b][r];
        st<true>(P.qout + idx, __builtin_fma(P.a, qs, P.b * __builtin_fma(P.dt, dq, qq)));
      }
    }
  }
}

template <bool F32, int NW>
void launch_mfma(const StageKernel &k, const StageParams &P, const HostOps &H, const double *,
                 hipStream_t st) {
  hipLaunchKernelGGL((mfma16_kernel<F32, NW>), dim3((unsigned)k.grid), dim3(64 * NW), k.lds, st,
                     P, H.Drow, H.lfac);
}

template <bool F32, int NW>
StageKernel make_mfma() {
  StageKernel k;
  k.kernel = (const void *)mfma16_kernel<F32, NW>;
  k.launch = launch_mfma<F32, NW>;
  k.nt = 64 * NW;
  k.lds = kMfmaLds;
  k.epb = 1;
  k.pencil = true;
  return k;
}

}  // namespace

StageKernel pick_mfma(int Nq, bool mixed, bool wide) {
  if (Nq != kMq) return StageKernel{};
  if (mixed) return wide ? make_mfma<true, 8>() : make_mfma<true, 4>();
  return wide ? make_mfma<false, 8>() : make_mfma<false, 4>();
}

}  // namespace dgk
