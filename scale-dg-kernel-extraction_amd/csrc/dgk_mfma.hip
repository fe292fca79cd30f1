// Matrix-core form of the fused RK stage for Nq = 16 (p = 15) with uniform
// velocity: the three directional contractions run as 16x16x4 MFMA products
// of the 1D operator with the element's nodal block, in FP64
// (v_mfma_f64_16x16x4_f64) or, for DGK_FLAG_MIXED, in FP32
// (v_mfma_f32_16x16x4_f32) on each line's remainder after its degree-7
// interpolant through the nodes 0, 2, 4, 6, 9, 11, 13, 15:
//   D q = D (q - P q) + (D P) q,
// with P the interpolation onto those nodes.  The interpolant P q and the
// correction (D P) q are FP64 MFMA products with K = 8 (the eight nodes'
// values against tabulated basis values and derivatives); only the
// remainder q - P q, small where the field is smooth, is rounded to FP32 and
// contracted with the FP32 operator, so the FP32 rounding scales with the
// remainder, not with the line.  Face fluxes, the factored lift, the residual
// sum and the RK update are FP64.
//
// One element per workgroup of NW = 4 or 8 wavefronts.  Wavefront w, lane
// L = 16 g + c holds in registers the nodes (i, j, k) = (c, BW w + b, k_r)
// for its BW = 16 / NW blocks b and the accumulator rows r = 0..3, with
// k_r = g + 4r for the FP64
// product and 4g + r for the FP32 one (the rows each MFMA form returns to a
// lane).  The block q(Np) is staged once in LDS (FP64); the x and z
// contractions read it as the MFMA operand of their k-by-i products and land
// directly in that register layout, both accumulated into the same registers
// (the scale factors Es_d u_d folded into the operator); the y contraction's
// k-by-j product goes back through LDS.  Three barriers per element.
//
// MFMA operand layout of the 16x16x4 forms (checked on the device): lane L
// supplies A[m = L % 16][kk = L / 16] and B[kk = L / 16][n = L % 16]; the
// result D[m][n] sits in lane n + 16 g as D[g + 4r][n] (FP64) or
// D[4g + r][n] (FP32), r = 0..3.  An FP64 product whose result joins an FP32
// one takes its A rows in the order tau(m) = 4 (m % 4) + m / 4, so that its
// row g + 4r carries row 4g + r.
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
// Interpolation node f of the mixed form's line fit (P above): 0, 2, 4, 6,
// 9, 11, 13, 15, the nodes on which mfma_fit_tables tabulates the basis.
__device__ __forceinline__ int fit_node(int f) { return f < 4 ? 2 * f : 2 * f + 1; }

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ d4 mma64(double a, double b, d4 acc) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
}
__device__ __forceinline__ f4 mma32(float a, float b, f4 acc) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
}

// No occupancy floor in the launch bounds: the four-wavefront FP64 form
// spills when held to three wavefronts per SIMD (47 % slower), so the
// compiler sizes the registers and the occupancy follows.
template <bool F32, int NW>
__global__ void __launch_bounds__(64 * NW, 1)
    mfma16_kernel(const StageParams P, const double *__restrict__ Drow,
                  const double *__restrict__ lfac, const double *__restrict__ mfit) {
  constexpr int BW = kMq / NW;  // blocks (j, or i for y) per wavefront
  extern __shared__ double smem[];
  double *Q = smem;  // q, then the y contraction
  double *Y = smem;
  double *F = smem + kQSlots;  // six face-flux planes of 256
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
    for (int r = 0; r < 4; ++r) Q[qpos(c, BW * w + b, krow(r))] = q[b][r];

  // operator operands: D(c, 4s + g), scaled by Es_d u_d of each direction
  using Op = typename std::conditional<F32, float, double>::type;
  const double sx = P.es[0] * cu, sy = P.es[1] * cv, sz = P.es[2] * cw;
  Op ax[4], ay[4], az[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const double d = Drow[kMq * c + 4 * s + g];
    ax[s] = (Op)(sx * d);
    ay[s] = (Op)(sy * d);
    az[s] = (Op)(sz * d);
  }
  __syncthreads();

  // ---- contractions -------------------------------------------------------
  // x: out(i,j,k) = sum_l D(i,l) q(l,j,k): A[m=k][kk=l] = q(l,j,k),
  //    B[kk=l][n=i] = D(i,l); z: out = sum_l D(k,l) q(i,j,l): A = D(k,l),
  //    B[kk=l][n=i] = q(i,j,l); both k-by-i for j = BW w + b.
  // y: out(i,j,k) = sum_l D(j,l) q(i,l,k): A[m=k][kk=l] = q(i,l,k),
  //    B[kk=l][n=j] = D(j,l), k-by-j for i = BW w + b.
  d4 acc[BW], accy[BW];  // x + z, and y, at (c, j or i = BW w + b, krow(r))
  if constexpr (!F32) {
#pragma unroll
    for (int b = 0; b < BW; ++b) {
      acc[b] = d4{0.0, 0.0, 0.0, 0.0};
      accy[b] = d4{0.0, 0.0, 0.0, 0.0};
      const int jb = BW * w + b;
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const int l = 4 * s + g;
        acc[b] = mma64(Q[qpos(l, jb, c)], ax[s], acc[b]);
        acc[b] = mma64(az[s], Q[qpos(c, jb, l)], acc[b]);
        accy[b] = mma64(Q[qpos(jb, l, c)], ay[s], accy[b]);
      }
    }
  } else {
    // Line fit: lane (c, g) supplies, for the eight-node chunk h, the basis
    // value l_f(x_c) (A of P q, f = 4h + g) and the derivative G(c, f) =
    // (D P)(c, f) (B of the x and y corrections; for z, whose correction
    // has the derivative as A, at row tau(c)).
    const int tc = 4 * (c & 3) + (c >> 2);
    double la[2], gd[2], gz[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      la[h] = mfit[8 * c + 4 * h + g];
      gd[h] = mfit[128 + 8 * c + 4 * h + g];
      gz[h] = sz * mfit[128 + 8 * tc + 4 * h + g];
    }
#pragma unroll
    for (int b = 0; b < BW; ++b) {
      const int jb = BW * w + b;
      // P q of the lane's x-line (jb, c), y-line (jb, c) and z-line (c, jb)
      // at the nodes l = g + 4r: its A/B operands in the remainders' layout
      d4 px = {0.0, 0.0, 0.0, 0.0}, py = px, pz = px;
      double bz[2];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int fn = fit_node(4 * h + g);
        bz[h] = Q[qpos(c, jb, fn)];
        px = mma64(la[h], Q[qpos(fn, jb, c)], px);
        py = mma64(la[h], Q[qpos(jb, fn, c)], py);
        pz = mma64(la[h], bz[h], pz);
      }
      // FP32 contractions of the remainders
      f4 a32 = {0.0f, 0.0f, 0.0f, 0.0f}, y32 = a32;
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const int l = 4 * s + g;
        a32 = mma32((float)(Q[qpos(l, jb, c)] - px[s]), ax[s], a32);
        a32 = mma32(az[s], (float)(Q[qpos(c, jb, l)] - pz[s]), a32);
        y32 = mma32((float)(Q[qpos(jb, l, c)] - py[s]), ay[s], y32);
      }
      // FP64 corrections (D P) q in the FP32 rows 4g + r
      d4 cxz = {0.0, 0.0, 0.0, 0.0}, cy = cxz;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int fn = fit_node(4 * h + g);
        cxz = mma64(Q[qpos(fn, jb, tc)], sx * gd[h], cxz);
        cxz = mma64(gz[h], bz[h], cxz);
        cy = mma64(Q[qpos(jb, fn, tc)], sy * gd[h], cy);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        acc[b][r] = (double)a32[r] + cxz[r];
        accy[b][r] = (double)y32[r] + cy[r];
      }
    }
  }
  __syncthreads();  // every read of the staged q is done
  // the y contraction at (i = BW w + b, j = c, k_r), over the staged q
#pragma unroll
  for (int b = 0; b < BW; ++b)
#pragma unroll
    for (int r = 0; r < 4; ++r) Y[qpos(BW * w + b, c, krow(r))] = accy[b][r];
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
      const double dq = -((acc[b][r] + Y[qpos(c, j, k)]) + lift);
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
                     P, H.Drow, H.lfac, H.mfit);
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
