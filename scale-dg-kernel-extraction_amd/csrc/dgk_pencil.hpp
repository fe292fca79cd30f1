// Pencil form of the fused DG advection stage for Nq in {2, 4, 8}: one
// 64-lane wavefront per workgroup, Nq^2 lanes per element, each lane owning
// one line of Nq nodes of its element at a time.
//
// Lane (i,j) loads the z-line (i,j,0..Nq-1) of q,u,v,w,q0: for every k the
// wavefront reads Nq^2 consecutive doubles, so all streaming loads and stores
// are fully coalesced.
//   z phase  registers: f_z = q w along the z-line, the z sweep and both
//            z-face fluxes need no data from other lanes.
//   y phase  q and v staged in LDS are read back along y-lines (i,k): the y
//            sweep and the two y-face fluxes; D_y returns to LDS.
//   x phase  q and u along x-lines (j,k): the x sweep and the two x-face
//            fluxes; D_x returns to LDS.
//   final    z-lines again: D_x, D_y and the x/y face fluxes come from LDS,
//            the lift, the combine and the SSP-RK3 update finish in
//            registers, and the new state is stored coalesced.
// Every LDS position belongs to exactly one y-line and one x-line, so a line
// may overwrite the slots it has read without a barrier.  LDS buffers pad
// every x-line by one double (index n + n/Nq): z-line writes, y-line and
// x-line reads of a wavefront then fall in distinct banks (but for one pair).
// The 1D operator D is read one row at a time into SGPRs with scalar loads,
// so every sweep multiplies by SGPR operands.  All neighbour-face
// loads are issued with the element's own loads at the start.  Arithmetic
// order is that of stage_kernel (the reference's order in exact variants).
#pragma once

#include "dgk_kernels.hpp"

namespace dgk {

template <int NQ>
struct PencilShape {
  static_assert(NQ >= 2 && NQ <= 16, "pencil stage: Nq in 2..16");
  static constexpr int NP = NQ * NQ * NQ;
  static constexpr int NFP = NQ * NQ;
  static constexpr int EPW = NFP <= 64 ? 64 / NFP : 1;  // elements per wavefront (Nq 2, 4, 8)
  static constexpr int PAD = NP + NP / NQ;     // padded volume buffer
  static constexpr int FPAD = NFP + NFP / NQ;  // padded face buffer
  static constexpr int PER_ELEM = 3 * PAD + 4 * FPAD;
  static constexpr size_t LDS = sizeof(double) * EPW * PER_ELEM;
  // LDS of one element in the lean layout: two volume buffers (q and the
  // partial sum S) and the four x/y face-flux buffers.
  // With uniform velocity and Nq >= 10 the group kernel's lean stage reads
  // the x- and y-lines of q from memory as well (lean_qg): the partial sum S
  // is then the only volume buffer, the element needs 40 % less LDS and
  // the wavefronts hold q in no register.  Measured 3-11 % faster at
  // Nq = 10, 12, 13, 16; 5 % slower at Nq = 14, with per-node velocity and
  // at lower orders, and in the ensemble (whose occupancy is set by VGPRs).
  static constexpr bool lean_qg(bool cv) { return cv && NQ >= 10 && NQ != 14; }
  static constexpr int lean_elem(bool cv) { return (lean_qg(cv) ? 1 : 2) * PAD + 4 * FPAD; }
};

// Threads and LDS of a multi-wavefront pencil workgroup of epb elements:
// Nq^2 lanes each, rounded up to whole wavefronts (the lanes beyond
// epb*Nq^2 only take part in the barriers).
template <int NQ>
constexpr int pencil_group_threads(int epb) { return (epb * NQ * NQ + 63) / 64 * 64; }
template <int NQ>
constexpr size_t pencil_group_lds(int epb, bool cv) {
  return sizeof(double) * epb * PencilShape<NQ>::lean_elem(cv);
}

// Padded LDS position of node n: one pad slot per x-line of Nq nodes.
template <int NQ>
__device__ __forceinline__ int pidx(int n) { return n + n / NQ; }

// Scalar loads of NDW consecutive dwords (a row of D: 2 Nq dwords) into
// SGPRs, in one statement with its wait: the widest loads first, each
// destination early-clobber so that no load overwrites the base address a
// later one still reads.  `after` orders the statement after a VGPR value.
template <int NDW>
__device__ __forceinline__ void sgpr_load(const double *p, double after, int (&dw)[NDW]);

template <>
__device__ __forceinline__ void sgpr_load<4>(const double *p, double after, int (&dw)[4]) {
  typedef int t0 __attribute__((ext_vector_type(4)));
  t0 x0;
  asm volatile("s_load_dwordx4 %0, %1, 0x0\n\ts_waitcnt lgkmcnt(0)"
               : "=s"(x0)
               : "s"(p), "v"(after));
#pragma unroll
  for (int t = 0; t < 4; ++t) dw[0 + t] = x0[t];
}

template <>
__device__ __forceinline__ void sgpr_load<6>(const double *p, double after, int (&dw)[6]) {
  typedef int t0 __attribute__((ext_vector_type(4)));
  t0 x0;
  typedef int t1 __attribute__((ext_vector_type(2)));
  t1 x1;
  asm volatile("s_load_dwordx4 %0, %2, 0x0\n\ts_load_dwordx2 %1, %2, 0x10\n\ts_waitcnt lgkmcnt(0)"
               : "=&s"(x0), "=&s"(x1)
               : "s"(p), "v"(after));
#pragma unroll
  for (int t = 0; t < 4; ++t) dw[0 + t] = x0[t];
#pragma unroll
  for (int t = 0; t < 2; ++t) dw[4 + t] = x1[t];
}

template <>
__device__ __forceinline__ void sgpr_load<8>(const double *p, double after, int (&dw)[8]) {
  typedef int t0 __attribute__((ext_vector_type(8)));
  t0 x0;
  asm volatile("s_load_dwordx8 %0, %1, 0x0\n\ts_waitcnt lgkmcnt(0)"
               : "=s"(x0)
               : "s"(p), "v"(after));
#pragma unroll
  for (int t = 0; t < 8; ++t) dw[0 + t] = x0[t];
}

template <>
__device__ __forceinline__ void sgpr_load<10>(const double *p, double after, int (&dw)[10]) {
  typedef int t0 __attribute__((ext_vector_type(8)));
  t0 x0;
  typedef int t1 __attribute__((ext_vector_type(2)));
  t1 x1;
  asm volatile("s_load_dwordx8 %0, %2, 0x0\n\ts_load_dwordx2 %1, %2, 0x20\n\ts_waitcnt lgkmcnt(0)"
               : "=&s"(x0), "=&s"(x1)
               : "s"(p), "v"(after));
#pragma unroll
  for (int t = 0; t < 8; ++t) dw[0 + t] = x0[t];
#pragma unroll
  for (int t = 0; t < 2; ++t) dw[8 + t] = x1[t];
}

template <>
__device__ __forceinline__ void sgpr_load<12>(const double *p, double after, int (&dw)[12]) {
  typedef int t0 __attribute__((ext_vector_type(8)));
  t0 x0;
  typedef int t1 __attribute__((ext_vector_type(4)));
  t1 x1;
  asm volatile("s_load_dwordx8 %0, %2, 0x0\n\ts_load_dwordx4 %1, %2, 0x20\n\ts_waitcnt lgkmcnt(0)"
               : "=&s"(x0), "=&s"(x1)
               : "s"(p), "v"(after));
#pragma unroll
  for (int t = 0; t < 8; ++t) dw[0 + t] = x0[t];
#pragma unroll
  for (int t = 0; t < 4; ++t) dw[8 + t] = x1[t];
}

template <>
__device__ __forceinline__ void sgpr_load<14>(const double *p, double after, int (&dw)[14]) {
  typedef int t0 __attribute__((ext_vector_type(8)));
  t0 x0;
  typedef int t1 __attribute__((ext_vector_type(4)));
  t1 x1;
  typedef int t2 __attribute__((ext_vector_type(2)));
  t2 x2;
  asm volatile("s_load_dwordx8 %0, %3, 0x0\n\ts_load_dwordx4 %1, %3, 0x20\n\ts_load_dwordx2 %2, %3, 0x30\n\ts_waitcnt lgkmcnt(0)"
               : "=&s"(x0), "=&s"(x1), "=&s"(x2)
               : "s"(p), "v"(after));
#pragma unroll
  for (int t = 0; t < 8; ++t) dw[0 + t] = x0[t];
#pragma unroll
  for (int t = 0; t < 4; ++t) dw[8 + t] = x1[t];
#pragma unroll
  for (int t = 0; t < 2; ++t) dw[12 + t] = x2[t];
}

template <>
__device__ __forceinline__ void sgpr_load<16>(const double *p, double after, int (&dw)[16]) {
  typedef int t0 __attribute__((ext_vector_type(16)));
  t0 x0;
  asm volatile("s_load_dwordx16 %0, %1, 0x0\n\ts_waitcnt lgkmcnt(0)"
               : "=s"(x0)
               : "s"(p), "v"(after));
#pragma unroll
  for (int t = 0; t < 16; ++t) dw[0 + t] = x0[t];
}

template <>
__device__ __forceinline__ void sgpr_load<18>(const double *p, double after, int (&dw)[18]) {
  typedef int t0 __attribute__((ext_vector_type(16)));
  t0 x0;
  typedef int t1 __attribute__((ext_vector_type(2)));
  t1 x1;
  asm volatile("s_load_dwordx16 %0, %2, 0x0\n\ts_load_dwordx2 %1, %2, 0x40\n\ts_waitcnt lgkmcnt(0)"
               : "=&s"(x0), "=&s"(x1)
               : "s"(p), "v"(after));
#pragma unroll
  for (int t = 0; t < 16; ++t) dw[0 + t] = x0[t];
#pragma unroll
  for (int t = 0; t < 2; ++t) dw[16 + t] = x1[t];
}

template <>
__device__ __forceinline__ void sgpr_load<20>(const double *p, double after, int (&dw)[20]) {
  typedef int t0 __attribute__((ext_vector_type(16)));
  t0 x0;
  typedef int t1 __attribute__((ext_vector_type(4)));
  t1 x1;
  asm volatile("s_load_dwordx16 %0, %2, 0x0\n\ts_load_dwordx4 %1, %2, 0x40\n\ts_waitcnt lgkmcnt(0)"
               : "=&s"(x0), "=&s"(x1)
               : "s"(p), "v"(after));
#pragma unroll
  for (int t = 0; t < 16; ++t) dw[0 + t] = x0[t];
#pragma unroll
  for (int t = 0; t < 4; ++t) dw[16 + t] = x1[t];
}

template <>
__device__ __forceinline__ void sgpr_load<22>(const double *p, double after, int (&dw)[22]) {
  typedef int t0 __attribute__((ext_vector_type(16)));
  t0 x0;
  typedef int t1 __attribute__((ext_vector_type(4)));
  t1 x1;
  typedef int t2 __attribute__((ext_vector_type(2)));
  t2 x2;
  asm volatile("s_load_dwordx16 %0, %3, 0x0\n\ts_load_dwordx4 %1, %3, 0x40\n\ts_load_dwordx2 %2, %3, 0x50\n\ts_waitcnt lgkmcnt(0)"
               : "=&s"(x0), "=&s"(x1), "=&s"(x2)
               : "s"(p), "v"(after));
#pragma unroll
  for (int t = 0; t < 16; ++t) dw[0 + t] = x0[t];
#pragma unroll
  for (int t = 0; t < 4; ++t) dw[16 + t] = x1[t];
#pragma unroll
  for (int t = 0; t < 2; ++t) dw[20 + t] = x2[t];
}

template <>
__device__ __forceinline__ void sgpr_load<24>(const double *p, double after, int (&dw)[24]) {
  typedef int t0 __attribute__((ext_vector_type(16)));
  t0 x0;
  typedef int t1 __attribute__((ext_vector_type(8)));
  t1 x1;
  asm volatile("s_load_dwordx16 %0, %2, 0x0\n\ts_load_dwordx8 %1, %2, 0x40\n\ts_waitcnt lgkmcnt(0)"
               : "=&s"(x0), "=&s"(x1)
               : "s"(p), "v"(after));
#pragma unroll
  for (int t = 0; t < 16; ++t) dw[0 + t] = x0[t];
#pragma unroll
  for (int t = 0; t < 8; ++t) dw[16 + t] = x1[t];
}

template <>
__device__ __forceinline__ void sgpr_load<26>(const double *p, double after, int (&dw)[26]) {
  typedef int t0 __attribute__((ext_vector_type(16)));
  t0 x0;
  typedef int t1 __attribute__((ext_vector_type(8)));
  t1 x1;
  typedef int t2 __attribute__((ext_vector_type(2)));
  t2 x2;
  asm volatile("s_load_dwordx16 %0, %3, 0x0\n\ts_load_dwordx8 %1, %3, 0x40\n\ts_load_dwordx2 %2, %3, 0x60\n\ts_waitcnt lgkmcnt(0)"
               : "=&s"(x0), "=&s"(x1), "=&s"(x2)
               : "s"(p), "v"(after));
#pragma unroll
  for (int t = 0; t < 16; ++t) dw[0 + t] = x0[t];
#pragma unroll
  for (int t = 0; t < 8; ++t) dw[16 + t] = x1[t];
#pragma unroll
  for (int t = 0; t < 2; ++t) dw[24 + t] = x2[t];
}

template <>
__device__ __forceinline__ void sgpr_load<28>(const double *p, double after, int (&dw)[28]) {
  typedef int t0 __attribute__((ext_vector_type(16)));
  t0 x0;
  typedef int t1 __attribute__((ext_vector_type(8)));
  t1 x1;
  typedef int t2 __attribute__((ext_vector_type(4)));
  t2 x2;
  asm volatile("s_load_dwordx16 %0, %3, 0x0\n\ts_load_dwordx8 %1, %3, 0x40\n\ts_load_dwordx4 %2, %3, 0x60\n\ts_waitcnt lgkmcnt(0)"
               : "=&s"(x0), "=&s"(x1), "=&s"(x2)
               : "s"(p), "v"(after));
#pragma unroll
  for (int t = 0; t < 16; ++t) dw[0 + t] = x0[t];
#pragma unroll
  for (int t = 0; t < 8; ++t) dw[16 + t] = x1[t];
#pragma unroll
  for (int t = 0; t < 4; ++t) dw[24 + t] = x2[t];
}

template <>
__device__ __forceinline__ void sgpr_load<30>(const double *p, double after, int (&dw)[30]) {
  typedef int t0 __attribute__((ext_vector_type(16)));
  t0 x0;
  typedef int t1 __attribute__((ext_vector_type(8)));
  t1 x1;
  typedef int t2 __attribute__((ext_vector_type(4)));
  t2 x2;
  typedef int t3 __attribute__((ext_vector_type(2)));
  t3 x3;
  asm volatile("s_load_dwordx16 %0, %4, 0x0\n\ts_load_dwordx8 %1, %4, 0x40\n\ts_load_dwordx4 %2, %4, 0x60\n\ts_load_dwordx2 %3, %4, 0x70\n\ts_waitcnt lgkmcnt(0)"
               : "=&s"(x0), "=&s"(x1), "=&s"(x2), "=&s"(x3)
               : "s"(p), "v"(after));
#pragma unroll
  for (int t = 0; t < 16; ++t) dw[0 + t] = x0[t];
#pragma unroll
  for (int t = 0; t < 8; ++t) dw[16 + t] = x1[t];
#pragma unroll
  for (int t = 0; t < 4; ++t) dw[24 + t] = x2[t];
#pragma unroll
  for (int t = 0; t < 2; ++t) dw[28 + t] = x3[t];
}

template <>
__device__ __forceinline__ void sgpr_load<32>(const double *p, double after, int (&dw)[32]) {
  typedef int t0 __attribute__((ext_vector_type(16)));
  t0 x0;
  typedef int t1 __attribute__((ext_vector_type(16)));
  t1 x1;
  asm volatile("s_load_dwordx16 %0, %2, 0x0\n\ts_load_dwordx16 %1, %2, 0x40\n\ts_waitcnt lgkmcnt(0)"
               : "=&s"(x0), "=&s"(x1)
               : "s"(p), "v"(after));
#pragma unroll
  for (int t = 0; t < 16; ++t) dw[0 + t] = x0[t];
#pragma unroll
  for (int t = 0; t < 16; ++t) dw[16 + t] = x1[t];
}

// Row k of the 1D operator, D(k,0..NQ-1), from the row-major copy Drow into
// SGPRs.  The loads are written out (with their wait in the same statement)
// because left to itself the compiler loads all NQ^2 coefficients at the top
// of the kernel and spills them to VGPR lanes, which costs more VALU issue
// slots than the sweeps themselves.  `after` is a value the row must not be
// loaded before (the previous output): it keeps one row of SGPRs live at a
// time.
template <int NQ>
__device__ __forceinline__ void sgpr_row(const double *Drow, int k, double after,
                                         double (&r)[NQ]) {
  int dw[2 * NQ];
  sgpr_load<2 * NQ>(Drow + k * NQ, after, dw);
#pragma unroll
  for (int l = 0; l < NQ; ++l)
    r[l] = __builtin_bit_cast(double, (unsigned long long)(unsigned)dw[2 * l] |
                                          ((unsigned long long)(unsigned)dw[2 * l + 1] << 32));
}

// FP32 row k of the split operator: the 2*NQ dwords of a row hold D_hi(k,.)
// then D_lo(k,.) (the same row length in dwords as the FP64 operator).
template <int NQ>
__device__ __forceinline__ void sgpr_row_split(const double *Drow, int k, double after,
                                               float (&hi)[NQ], float (&lo)[NQ]) {
  double r[NQ];
  sgpr_row<NQ>(Drow, k, after, r);
#pragma unroll
  for (int l = 0; l < NQ; ++l) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, r[l]);
    const unsigned w0 = (unsigned)b, w1 = (unsigned)(b >> 32);
    // dword 2l and 2l+1 of the row
    const int d0 = 2 * l, d1 = 2 * l + 1;
    if (d0 < NQ) hi[d0] = __builtin_bit_cast(float, w0); else lo[d0 - NQ] = __builtin_bit_cast(float, w0);
    if (d1 < NQ) hi[d1] = __builtin_bit_cast(float, w1); else lo[d1 - NQ] = __builtin_bit_cast(float, w1);
  }
}

// Rows of D in flight during a sweep: row k is loaded once output
// k - kSweepDepth is done (kSweepDepth rows of SGPRs live).  A depth of 2,
// which overlaps a row's load with the previous output's arithmetic, measured
// no faster at p = 7, 11, 15 than 1.
constexpr int kSweepDepth = 1;

// A line of the mixed form whose FP64 remainder after the fit exceeds this
// fraction of the line's largest value is contracted in FP64.
constexpr double kMixedRough = 1.0 / 16.0;

// The line fit of the mixed form: the polynomial of degree DEG through the
// nodes 0, A, B and NQ-1 of a line (Newton basis 1, t, t(t-1),
// t(t-1)(t-t_A) in t = (x+1)/2) is differentiated exactly in FP64; only the
// remainder, zero at those nodes, goes through FP32.
template <int NQ>
struct LineFit {
  static constexpr int DEG = NQ >= 4 ? 3 : NQ - 1;
  static constexpr int A = NQ / 3, B = 2 * NQ / 3;
  static constexpr bool fit(int l) {
    return l == 0 || l == NQ - 1 || (DEG >= 2 && l == A) || (DEG >= 3 && l == B);
  }
};

__device__ __forceinline__ double dword_pair(float lo, float hi) {
  return __builtin_bit_cast(double, (unsigned long long)__builtin_bit_cast(unsigned, lo) |
                                        ((unsigned long long)__builtin_bit_cast(unsigned, hi) << 32));
}

// out[k] = sum_l D(k,l) f[l] in the reference's order, one SGPR row at a time.
// ORDER == kMixed: the FP32 form, Drow holding the split rows (D_hi(k,.),
// D_lo(k,.) as FP32) and after them the rows t, p2, p3 of the line fit
// (split_tables).  f = g + r with g the fit polynomial (LineFit): D g = g'
// exactly at the nodes (D differentiates polynomials of degree <= p
// exactly), evaluated in FP64 as (b + c2 p2' + c3 p3')/2; r, zero at the fit
// nodes and small against f where f is smooth across the element, is
// rounded to FP32 and contracted as D_hi r + D_lo r in two FP32 sums.  The
// FP32 rounding then scales with |r| instead of |f|: at 64^3, p=7 on the
// sine-mode field the contraction error drops ~1000x against hi/lo-split
// operands of the whole line; at p <= 3 the fit is the whole line and no
// FP32 arithmetic is left in the sweep.  Lines with a large remainder
// (kMixedRough) take the FP64 contraction.
template <int NQ, int ORDER, bool FMA, bool MAT_FIRST>
__device__ __forceinline__ void sweep(const double *Drow, const double (&f)[NQ],
                                      double (&out)[NQ]) {
  if constexpr (ORDER == kMixed) {
    using F = LineFit<NQ>;
    constexpr int A = F::A, B = F::B, DEG = F::DEG;
    const double f0 = f[0];
    const double b = f[NQ - 1] - f0;
    double rr[NQ];
    double inv2 = 0.0, inv3 = 0.0;
    {
      double t[NQ];
      sgpr_row<NQ>(Drow, NQ, f0, t);  // t_l; end slots: the Newton divisors
      inv2 = t[0];
      inv3 = t[NQ - 1];
#pragma unroll
      for (int l = 1; l < NQ - 1; ++l) rr[l] = __builtin_fma(-b, t[l], f[l] - f0);
    }
    double c2 = 0.0, c3 = 0.0;
    if constexpr (DEG >= 2) {
      c2 = rr[A] * inv2;
      double p2[NQ];
      sgpr_row<NQ>(Drow, NQ + 1, c2, p2);
#pragma unroll
      for (int l = 1; l < NQ - 1; ++l)
        if (l != A) rr[l] = __builtin_fma(-c2, p2[l], rr[l]);
    }
    if constexpr (DEG >= 3) {
      c3 = rr[B] * inv3;
      double p3[NQ];
      sgpr_row<NQ>(Drow, NQ + 2, c3, p3);
#pragma unroll
      for (int l = 1; l < NQ - 1; ++l)
        if (!F::fit(l)) rr[l] = __builtin_fma(-c3, p3[l], rr[l]);
    }
    // A line whose remainder is not small against the line itself (next to
    // a discontinuity, or a field not resolved by the element) is
    // contracted in FP64 instead: FP32 would round it at 2^-24 of the field.
    double rmax = 0.0, fmx = 0.0;
#pragma unroll
    for (int l = 0; l < NQ; ++l) {
      fmx = fmax(fmx, fabs(f[l]));
      if (!F::fit(l)) rmax = fmax(rmax, fabs(rr[l]));
    }
    const bool rough = rmax > kMixedRough * fmx;
    float r[NQ];
#pragma unroll
    for (int l = 0; l < NQ; ++l) r[l] = F::fit(l) ? 0.0f : (float)rr[l];
    const double half_b = 0.5 * b;
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
      float dh[NQ], dl[NQ];
      sgpr_row_split<NQ>(Drow, k, k >= kSweepDepth ? out[k - kSweepDepth] : f0, dh, dl);
      double g = half_b;
      if constexpr (DEG >= 2) g = __builtin_fma(c2, dword_pair(dh[0], dh[NQ - 1]), g);
      if constexpr (DEG >= 3) g = __builtin_fma(c3, dword_pair(dl[0], dl[NQ - 1]), g);
      if (__builtin_expect(rough, 0)) {
        // D = D_hi + D_lo to 2^-48 relative
        double acc = 0.0;
#pragma unroll
        for (int l = 1; l < NQ - 1; ++l) {
          if (F::fit(l)) continue;
          acc = __builtin_fma((double)dh[l] + (double)dl[l], rr[l], acc);
        }
        out[k] = acc + g;
      } else {
        float acc = 0.0f, cor = 0.0f;
#pragma unroll
        for (int l = 1; l < NQ - 1; ++l) {
          if (F::fit(l)) continue;
          acc = __builtin_fmaf(dh[l], r[l], acc);
          cor = __builtin_fmaf(dl[l], r[l], cor);
        }
        out[k] = ((double)acc + (double)cor) + g;
      }
    }
    return;
  }
#pragma unroll
  for (int k = 0; k < NQ; ++k) {
    double r[NQ];
    sgpr_row<NQ>(Drow, k, k >= kSweepDepth ? out[k - kSweepDepth] : f[0], r);
    // x: Mat(i,l)*in(l); y,z: in(l)*Mat_tr(l,j) -- products commute
    out[k] = contract<NQ, ORDER, FMA>([&](int l) { return r[l]; }, [&](int l) { return f[l]; });
  }
}

// Tuning variants of the pencil kernel (bit set):
//   kPencilNtStore      write q, q0 and dqdt with non-temporal stores (no
//                       reuse within a launch)
//   kPencilLinearOrder  linear element order instead of the tiled order on
//                       structured meshes (a launch parameter, not a
//                       template argument)
//   kPencilWide         Nq 2, 4, 8: the group kernel's lean layout (one
//                       wavefront); Nq 3, 5, 6, 7, 9-12: more elements per
//                       workgroup (Nq 13-16 have one shape)
constexpr int kPencilNtStore = 1;
constexpr int kPencilLinearOrder = 2;
constexpr int kPencilWide = 4;
constexpr int kPencilVariants = 8;

// Lean stage: request the x-lines of u, the y-lines of v and q0 one phase
// ahead of their use (see pencil_compute).
constexpr bool kLeanEarlyLoads = true;

// Wavefront priority around a phase's loads (PRIO): a wavefront that is
// about to request its next lines is issued ahead of the others on its SIMD,
// so that its loads leave before their arithmetic.  Measured on the
// persistent ensemble kernel (161 -> 168 GDOF-updates/s on the 64-instance
// workload); the one-element-per-workgroup stages ran 2-7 % slower with it.
template <bool PRIO>
__device__ __forceinline__ void raise_prio() {
  if constexpr (PRIO) __builtin_amdgcn_s_setprio(2);
}
template <bool PRIO>
__device__ __forceinline__ void lower_prio() {
  if constexpr (PRIO) __builtin_amdgcn_s_setprio(0);
}

template <bool NT>
__device__ __forceinline__ void st(double *p, double v) {
  if constexpr (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

// CV: constant velocity (P.cvel): only q and q0 stream through HBM; the
// products q*u etc. and the face velocities are those of uniform u,v,w
// fields, so the results are the same bits.
// The global loads of one lane of the pencil stage: its own z-line
// (i,j) = (la,lb) of q, u, v, w (and q0) and the six neighbour face traces.
template <int NQ>
struct PencilLoads {
  double q[NQ], u[NQ], v[NQ], w[NQ], q0[NQ];
  double qP0, qP1, qP2, qP3, qP4, qP5, vP0, uP1, vP2, uP3, wP4, wP5;
};

// Q0: also load q0 (stages 2 and 3 need it).  UV: also load the z-lines of
// u and v (the lean layout reads their x- and y-lines itself).  QZ: load the
// z-line of q (the lean layout with q lines from memory reads it at the
// end).  eb is a valid element.
template <int NQ, bool CV, bool Q0, bool UV, class PT, bool QZ = true>
__device__ __forceinline__ void pencil_load(PT &P, long long eb, int pl, PencilLoads<NQ> &L) {
  constexpr int NP = NQ * NQ * NQ, NFP = NQ * NQ;
  const int la = pl % NQ, lb = pl / NQ;
  const long long base = eb * NP + pl;
  const bool need_q0 = P.mode == kModeStageN;
#pragma unroll
  for (int k = 0; k < NQ; ++k) {
    if constexpr (QZ) L.q[k] = P.qin[base + NFP * k];
    if constexpr (CV) {
      L.u[k] = P.cvel[0];
      L.v[k] = P.cvel[1];
      L.w[k] = P.cvel[2];
    } else {
      if constexpr (UV) {
        L.u[k] = P.u[base + NFP * k];
        L.v[k] = P.v[base + NFP * k];
      }
      L.w[k] = P.w[base + NFP * k];
    }
    if constexpr (Q0) L.q0[k] = need_q0 ? ld_stream(P.q0 + base + NFP * k) : 0.0;
  }
  const int32_t *nb = P.nbr + eb * 6;
  // z faces at (i,j): -z (face 4) sees the neighbour's k = NQ-1, +z (face 5)
  // its k = 0
  const long long a4 = (long long)nb[4] * NP + pl + NFP * (NQ - 1);
  const long long a5 = (long long)nb[5] * NP + pl;
  // y faces at (i,k) = (la,lb): -y (face 0) sees j = NQ-1, +y (face 2) j = 0
  const long long a0 = (long long)nb[0] * NP + la + NQ * (NQ - 1) + NFP * lb;
  const long long a2 = (long long)nb[2] * NP + la + NFP * lb;
  // x faces at (j,k) = (la,lb): -x (face 3) sees i = NQ-1, +x (face 1) i = 0
  const long long a3 = (long long)nb[3] * NP + (NQ - 1) + NQ * pl;
  const long long a1 = (long long)nb[1] * NP + NQ * pl;
  L.qP4 = P.qin[a4];
  L.qP5 = P.qin[a5];
  L.qP0 = P.qin[a0];
  L.qP2 = P.qin[a2];
  L.qP3 = P.qin[a3];
  L.qP1 = P.qin[a1];
  L.wP4 = CV ? P.cvel[2] : P.w[a4];
  L.wP5 = CV ? P.cvel[2] : P.w[a5];
  L.vP0 = CV ? P.cvel[1] : P.v[a0];
  L.vP2 = CV ? P.cvel[1] : P.v[a2];
  L.uP3 = CV ? P.cvel[0] : P.u[a3];
  L.uP1 = CV ? P.cvel[0] : P.u[a1];
}

// The stage of one element from its loaded lines.  LATE_Q0: q0 is not in L
// and is read here, at the start (non-LEAN) or in the final phase (LEAN).
template <int NQ, int ORDER, bool FMA, bool FACT_LIFT, bool NT, bool CV, bool LEAN, bool LATE_Q0,
          class PT, bool QG = false, bool PRIO = false>
__device__ __forceinline__ void pencil_compute(PT &P, const double *__restrict__ Drow,
                                               const double *__restrict__ Lift,
                                               const double *__restrict__ lfac, double *blk,
                                               int pl, long long e, long long eb, bool act,
                                               PencilLoads<NQ> &L) {
  // Lift: the delta-packed copy described at the lift loads below
  using S = PencilShape<NQ>;
  constexpr int NP = S::NP, NFP = S::NFP, PAD = S::PAD, FPAD = S::FPAD;
  // QG (lean only): the x- and y-lines of q are read from memory, not LDS
  static_assert(!QG || LEAN, "q lines from memory need the lean layout");
  constexpr int NVOL = LEAN ? (QG ? 1 : 2) : 3;
  const int la = pl % NQ, lb = pl / NQ;
  // Padded LDS positions of this lane's lines.  For Nq not a power of two
  // pidx's division would be evaluated at every access, so the positions
  // take their closed forms (l, i, k < NQ):
  //   z-line (la,lb,k): pidx(pl + NFP k) = zb + (NFP + NQ) k
  //   x-line (l,la,lb): pidx(l + NQ pl) = xb + l
  //   y-line (la,l,lb): pidx(la + NQ l + NFP lb) = yb + (NQ + 1) l
  //   face slots: pidx(pl) = zb, pidx(i + NQ k) = i + (NQ + 1) k
  // For a power of two the division is a shift and pidx itself needs fewer
  // live registers, which counts where VGPRs set the occupancy: the
  // one-wavefront stage and the lean stage reading q from memory (Nq = 16,
  // uniform-velocity FMA form: 162 against 171 VGPRs, three wavefronts per
  // SIMD instead of two).  The lean stage with q in LDS is held to two
  // elements per CU at Nq = 16 by its LDS anyway, and there the closed forms
  // let the compiler pair neighbouring reads into ds_read2_b64 (p=15,
  // per-node velocity: 124 instead of 196 LDS instructions per wavefront,
  // the stage 3-4 % faster).
  constexpr bool kShift = (NQ & (NQ - 1)) == 0 && (QG || !LEAN);
  const int zb = pl + lb, xb = (NQ + 1) * pl, yb = la + (NFP + NQ) * lb;
  auto pz = [&](int k) { return kShift ? pidx<NQ>(pl + NFP * k) : zb + (NFP + NQ) * k; };
  auto px = [&](int l) { return kShift ? pidx<NQ>(l + NQ * pl) : xb + l; };
  auto py = [&](int l) { return kShift ? pidx<NQ>(la + NQ * l + NFP * lb) : yb + (NQ + 1) * l; };
  auto pf = [&](int i, int k) { return kShift ? pidx<NQ>(i + NQ * k) : i + (NQ + 1) * k; };
  const int ps = kShift ? pidx<NQ>(pl) : zb;
  double *Bq = blk;
  double *Bu = Bq + PAD;  // u, then D_x (LEAN: the partial sum S)
  double *Bv = Bu + PAD;  // v, then D_y
  double *BS = QG ? blk : Bu;  // LEAN: Es1 D_x, then + Es2 D_y
  double *Bf = Bq + NVOL * PAD;  // fluxes of faces -y, +y, -x, +x
  const bool valid = e >= 0;
  const long long ebase = eb * NP;
  const long long base = ebase + pl;
  const bool need_q0 = P.mode == kModeStageN;
  double (&q)[NQ] = L.q;
  double (&u)[NQ] = L.u;
  double (&v)[NQ] = L.v;
  double (&w)[NQ] = L.w;
  double (&q0)[NQ] = L.q0;
  if constexpr (LATE_Q0 && !LEAN) {
#pragma unroll
    for (int k = 0; k < NQ; ++k) q0[k] = need_q0 ? ld_stream(P.q0 + base + NFP * k) : 0.0;
  }
  const double qP4 = L.qP4, qP5 = L.qP5, qP0 = L.qP0, qP2 = L.qP2, qP3 = L.qP3, qP1 = L.qP1;
  const double wP4 = L.wP4, wP5 = L.wP5, vP0 = L.vP0, vP2 = L.vP2, uP3 = L.uP3, uP1 = L.uP1;

  // ---- z phase (registers; LEAN: in the final phase) ---------------------
  // LEAN keeps neither q, q0 nor D_z in registers through the x and y
  // phases: q is read back from LDS and q0 from memory at the end.
  double dz[NQ];
  double flux4 = 0.0, flux5 = 0.0;
  auto z_phase = [&](const double (&qz)[NQ]) {
    double fz[NQ];
#pragma unroll
    for (int k = 0; k < NQ; ++k) fz[k] = qz[k] * w[k];
    sweep<NQ, ORDER, FMA, false>(Drow, fz, dz);
    flux4 = upwind_flux(qz[0], qP4, -w[0], -wP4, P.fs[4]);
    flux5 = upwind_flux(qz[NQ - 1], qP5, w[NQ - 1], wP5, P.fs[5]);
  };
  // LEAN: every line a phase reads from memory (the x-lines of u, the y-lines
  // of v, q0) is requested before the barrier that ends the previous phase,
  // so that its latency overlaps that wait instead of heading the phase.
  constexpr bool kEarly = LEAN && kLeanEarlyLoads;
  double uxe[NQ], vye[NQ];
  auto load_ux = [&] {
#pragma unroll
    for (int l = 0; l < NQ; ++l) uxe[l] = CV ? P.cvel[0] : P.u[ebase + l + NQ * pl];
  };
  auto load_vy = [&] {
#pragma unroll
    for (int l = 0; l < NQ; ++l) vye[l] = CV ? P.cvel[1] : P.v[ebase + la + NQ * l + NFP * lb];
  };
  auto load_q0 = [&] {
#pragma unroll
    for (int k = 0; k < NQ; ++k) q0[k] = need_q0 ? ld_stream(P.q0 + base + NFP * k) : 0.0;
  };
  // Dense lift, delta-packed (pack_lift_deltas): Lift(i,j,k,f) is the double
  // whose bits are those of the base of (f, normal index) plus the unsigned
  // byte f of this line's 8-byte word for node k.  The words of nodes 2r and
  // 2r+1 of a line are adjacent, words [2r NFP + 2 pl] and [.. + 1] (lanes
  // contiguous, one 16-byte load per pair), and an odd Nq's last node has
  // word [(NQ-1) NFP + pl].  The bases of the four side faces are fixed along
  // the line; those of -z and +z follow k and are the same for every lane.
  // The words of the lane's Nq nodes are requested together ahead of the
  // final phase: loaded inside its loop, between the stores, each node
  // waited on its own cache round trip.
  const long long *LB = reinterpret_cast<const long long *>(Lift);
  const uint2 *LW = reinterpret_cast<const uint2 *>(Lift + 6 * NQ);
  uint2 lw[FACT_LIFT ? 1 : NQ];
  auto load_lw = [&] {
    if constexpr (!FACT_LIFT) {
#pragma unroll
      for (int r = 0; r < NQ / 2; ++r) {
        const uint4 two = *reinterpret_cast<const uint4 *>(LW + 2 * r * NFP + 2 * pl);
        lw[2 * r] = make_uint2(two.x, two.y);
        lw[2 * r + 1] = make_uint2(two.z, two.w);
      }
      if constexpr (NQ % 2 == 1) lw[NQ - 1] = LW[(NQ - 1) * NFP + pl];
    }
  };
  // Factored lift: the coefficients of the four side faces are constant
  // along the lane's line, l_-(j), l_+(i), l_+(j), l_-(i) (non-lean stage),
  // read as the pairs (l_-(n), l_+(n)) that follow [l_-, l_+] in lfac.
  double lside[4] = {0.0, 0.0, 0.0, 0.0};
  auto load_lside = [&] {
    if constexpr (FACT_LIFT) {
      const double2 *pairs = reinterpret_cast<const double2 *>(lfac + 2 * NQ);
      const double2 pj = pairs[lb], pi = pairs[la];
      lside[0] = pj.x;
      lside[1] = pi.y;
      lside[2] = pj.y;
      lside[3] = pi.x;
    }
  };
  // Non-lean stage at Nq >= 4: the final phase's lift data is requested
  // here, behind the element's own loads, and the z phase runs while it
  // arrives (at Nq = 2 it measured 1 % slower than loading it at the end).
  constexpr bool kLiftAhead = !LEAN && NQ >= 4;
  if constexpr (!LEAN) {
    if constexpr (kLiftAhead) {
      load_lw();
      load_lside();
    }
    z_phase(q);
  }
  if constexpr (kEarly) {
    raise_prio<PRIO>();
    load_ux();
    lower_prio<PRIO>();
  }
  if constexpr (!QG) {
    if (act) {
#pragma unroll
      for (int k = 0; k < NQ; ++k) {
        const int s = pz(k);
        Bq[s] = q[k];
        if constexpr (!CV && !LEAN) {
          Bu[s] = u[k];
          Bv[s] = v[k];
        }
      }
    }
    __syncthreads();
  }
  const double e1s = P.es[0], e2s = P.es[1], e3s = P.es[2];

  if constexpr (LEAN) {
    // ---- x phase: line (j,k) = (la,lb); S = Es1 D_x ------------------------
    // The x-line of u is read from memory (Nq consecutive values per lane).
    {
      double qx[NQ], fx[NQ], dx[NQ];
      if constexpr (!kEarly) load_ux();
      const double (&ux)[NQ] = uxe;
#pragma unroll
      for (int l = 0; l < NQ; ++l) {
        qx[l] = QG ? P.qin[ebase + l + NQ * pl] : Bq[px(l)];
        fx[l] = qx[l] * ux[l];
      }
      sweep<NQ, ORDER, FMA, true>(Drow, fx, dx);
      const double f3 = upwind_flux(qx[0], qP3, -ux[0], -uP3, P.fs[3]);
      const double f1 = upwind_flux(qx[NQ - 1], qP1, ux[NQ - 1], uP1, P.fs[1]);
      if (act) {
        Bf[2 * FPAD + ps] = f3;
        Bf[3 * FPAD + ps] = f1;
#pragma unroll
        for (int i = 0; i < NQ; ++i) BS[px(i)] = e1s * dx[i];
      }
    }
    if constexpr (kEarly) {
      raise_prio<PRIO>();
      load_vy();
      lower_prio<PRIO>();
    }
    __syncthreads();
    // ---- y phase: line (i,k) = (la,lb); S += Es2 D_y -----------------------
    // The y-line of v is read from memory.
    {
      double qy[NQ], fy[NQ], dy[NQ];
      if constexpr (!kEarly) load_vy();
      const double (&vy)[NQ] = vye;
#pragma unroll
      for (int l = 0; l < NQ; ++l) {
        qy[l] = QG ? P.qin[ebase + la + NQ * l + NFP * lb] : Bq[py(l)];
        fy[l] = qy[l] * vy[l];
      }
      sweep<NQ, ORDER, FMA, false>(Drow, fy, dy);
      const double f0 = upwind_flux(qy[0], qP0, -vy[0], -vP0, P.fs[0]);
      const double f2 = upwind_flux(qy[NQ - 1], qP2, vy[NQ - 1], vP2, P.fs[2]);
      if (act) {
        Bf[0 * FPAD + ps] = f0;
        Bf[1 * FPAD + ps] = f2;
#pragma unroll
        for (int j = 0; j < NQ; ++j) {
          const int s = py(j);
          BS[s] = madd<FMA>(e2s, dy[j], BS[s]);
        }
      }
    }
    if constexpr (kEarly) {
      raise_prio<PRIO>();
      load_q0();
      load_lw();
      lower_prio<PRIO>();
    }
    __syncthreads();
  } else {
    // The lift data requested with the element's loads is taken into
    // registers here, a phase ahead of its use.  With no use before the final
    // phase the compiler moves those loads down to it, past the return of
    // idle lanes, and each element then waits there on their round trip
    // (64^3 p=7, uniform-velocity FMA stage: 624 -> 602 us).
    if constexpr (kLiftAhead && FACT_LIFT) {
      asm volatile("" ::"v"(lside[0]), "v"(lside[1]), "v"(lside[2]), "v"(lside[3]));
    } else if constexpr (kLiftAhead) {
#pragma unroll
      for (int k = 0; k < NQ; ++k) asm volatile("" ::"v"(lw[k].x), "v"(lw[k].y));
    }
    // ---- y phase: line (i,k) = (la,lb) -------------------------------------
    {
      double qy[NQ], vy[NQ], fy[NQ], dy[NQ];
#pragma unroll
      for (int l = 0; l < NQ; ++l) {
        const int s = py(l);
        qy[l] = Bq[s];
        vy[l] = CV ? P.cvel[1] : Bv[s];
        fy[l] = qy[l] * vy[l];
      }
      sweep<NQ, ORDER, FMA, false>(Drow, fy, dy);
      const double f0 = upwind_flux(qy[0], qP0, -vy[0], -vP0, P.fs[0]);
      const double f2 = upwind_flux(qy[NQ - 1], qP2, vy[NQ - 1], vP2, P.fs[2]);
      if (act) {
        Bf[0 * FPAD + ps] = f0;
        Bf[1 * FPAD + ps] = f2;
#pragma unroll
        for (int j = 0; j < NQ; ++j) Bv[py(j)] = dy[j];
      }
    }

    // ---- x phase: line (j,k) = (la,lb) -------------------------------------
    {
      double qx[NQ], ux[NQ], fx[NQ], dx[NQ];
#pragma unroll
      for (int l = 0; l < NQ; ++l) {
        const int s = px(l);
        qx[l] = Bq[s];
        ux[l] = CV ? P.cvel[0] : Bu[s];
        fx[l] = qx[l] * ux[l];
      }
      sweep<NQ, ORDER, FMA, true>(Drow, fx, dx);
      const double f3 = upwind_flux(qx[0], qP3, -ux[0], -uP3, P.fs[3]);
      const double f1 = upwind_flux(qx[NQ - 1], qP1, ux[NQ - 1], uP1, P.fs[1]);
      if (act) {
        Bf[2 * FPAD + ps] = f3;
        Bf[3 * FPAD + ps] = f1;
#pragma unroll
        for (int i = 0; i < NQ; ++i) Bu[px(i)] = dx[i];
      }
    }
    __syncthreads();
  }

  // ---- final z phase: lift, combine, RK ----------------------------------
  if (!valid) return;
  const int i = la, j = lb;
  if constexpr (!kLiftAhead && !kEarly) load_lw();
  if constexpr (!LEAN && !kLiftAhead) load_lside();
  long long b0 = 0, b1 = 0, b2 = 0, b3 = 0;
  if constexpr (!FACT_LIFT) {
    b0 = LB[0 * NQ + j];
    b1 = LB[1 * NQ + i];
    b2 = LB[2 * NQ + j];
    b3 = LB[3 * NQ + i];
  }
  auto delta = [](const uint2 &wd, int f) -> long long {  // byte f of a word
    const unsigned h = f < 4 ? wd.x : wd.y;
    return (long long)((h >> (8 * (f % 4))) & 0xffu);
  };
  if constexpr (LEAN) {
    if constexpr (!kEarly) load_q0();
#pragma unroll
    for (int k = 0; k < NQ; ++k) q[k] = QG ? P.qin[base + NFP * k] : Bq[pz(k)];
    z_phase(q);
  }
  // The loop has no branch: the tendency launch (dqdt) and the RK stages (q)
  // differ in a uniform choice of store target and value.  With the mode
  // tested inside it every node was a basic block of its own and waited on
  // its own scalar loads (the -z/+z lift coefficients of its k) instead of
  // having them requested together.
  const bool tend = P.mode == kModeTend;
  double *const dst = tend ? P.dqdt : P.qout;
  const bool stage1 = P.mode == kModeStage1;
#pragma unroll
  for (int k = 0; k < NQ; ++k) {
    const int s = pz(k);
    const double f0 = Bf[0 * FPAD + pf(i, k)];  // in(i,k,1)
    const double f2 = Bf[1 * FPAD + pf(i, k)];  // in(i,k,3)
    const double f3 = Bf[2 * FPAD + pf(j, k)];  // in(j,k,4)
    const double f1 = Bf[3 * FPAD + pf(j, k)];  // in(j,k,2)
    double L[6];
    static_assert(ORDER != kMixed || FACT_LIFT, "the mixed form lifts with the 1D vectors");
    if constexpr (FACT_LIFT) {
      // l_-(normal index) for -y,-x,-z; l_+ for +x,+y,+z; lfac = [l_-, l_+]
      // (lean: read here; requesting them a phase ahead measured no faster)
      L[0] = LEAN ? lfac[j] : lside[0];
      L[1] = LEAN ? lfac[NQ + i] : lside[1];
      L[2] = LEAN ? lfac[NQ + j] : lside[2];
      L[3] = LEAN ? lfac[i] : lside[3];
      L[4] = lfac[k];
      L[5] = lfac[NQ + k];
    } else {
      const uint2 wd = lw[k];
      L[0] = __longlong_as_double(b0 + delta(wd, 0));
      L[1] = __longlong_as_double(b1 + delta(wd, 1));
      L[2] = __longlong_as_double(b2 + delta(wd, 2));
      L[3] = __longlong_as_double(b3 + delta(wd, 3));
      L[4] = __longlong_as_double(LB[4 * NQ + k] + delta(wd, 4));
      L[5] = __longlong_as_double(LB[5 * NQ + k] + delta(wd, 5));
    }
    // tensorprod_Lift_hexahedral order (mod_dg_optr_kernel.f90:277-283).
    // The mixed form lifts in FP64 too (fused): a lift term has no
    // cancellation to take out, so FP32 would round it at 2^-24 of the
    // face-flux jump, which dominates the error next to a discontinuity,
    // and six FP64 FMAs per node cost fewer instructions than a split FP32
    // form.
    constexpr bool kFuseLift = FMA || ORDER == kMixed;
    double lift = L[0] * f0;
    lift = madd<kFuseLift>(L[1], f1, lift);
    lift = madd<kFuseLift>(L[2], f2, lift);
    lift = madd<kFuseLift>(L[3], f3, lift);
    lift = madd<kFuseLift>(L[4], flux4, lift);
    lift = madd<kFuseLift>(L[5], flux5, lift);
    // dqdt = -(Es1 Dx + Es2 Dy + Es3 Dz + Lift) (mod_advect3d_eq.f90:194-198)
    double sum;
    if constexpr (LEAN) {
      sum = BS[s];
    } else {
      sum = e1s * Bu[s];
      sum = madd<FMA>(e2s, Bv[s], sum);
    }
    sum = madd<FMA>(e3s, dz[k], sum);
    const double dq = -(sum + lift);
    const long long idx = base + NFP * k;
    const double qq = q[k];
    const double qs = stage1 ? qq : q0[k];
    double qn;
    if constexpr (FMA) qn = __builtin_fma(P.a, qs, P.b * __builtin_fma(P.dt, dq, qq));
    else qn = P.a * qs + P.b * (qq + P.dt * dq);
    st<NT>(dst + idx, tend ? dq : qn);
  }
}

// One element's lines (or EPW elements of one wavefront) of the pencil stage.
// e is this lane's element or -1 (idle lane), e_any any element of the same
// workgroup (for the address math of idle lanes), blk the LDS block of this
// lane's element, pl the lane's line index a + NQ*b.  act = false marks a
// lane beyond the workgroup's elements: it writes no LDS (blk may then be
// any block).  Every lane of the workgroup must call it (it holds two
// workgroup barriers, three when LEAN).
// LEAN: the x phase runs first and leaves Es1*D_x in a buffer S that the y
// phase turns into Es1*D_x + Es2*D_y (the reference's summation order); the
// x- and y-lines of u and v come straight from memory, so two volume buffers
// (q and S) suffice instead of three, whatever the velocity.
// PT: the parameter block's type, StageParams or a StageParams in the
// constant address space (read on demand with scalar loads).  PRIO: each
// phase's loads are issued at raised wavefront priority (raise_prio).
template <int NQ, int ORDER, bool FMA, bool FACT_LIFT, bool NT, bool CV, bool LEAN = false,
          class PT = const StageParams, bool QG = false, bool PRIO = false>
__device__ __forceinline__ void pencil_element(PT &P, const double *__restrict__ Drow,
                                               const double *__restrict__ Lift,
                                               const double *__restrict__ lfac, double *blk,
                                               int pl, long long e, long long e_any,
                                               bool act = true) {
  const long long eb = e >= 0 ? e : e_any;  // a valid element for address math
  PencilLoads<NQ> L;
  // LEAN reads q0 in the final phase
  raise_prio<PRIO>();
  pencil_load<NQ, CV, !LEAN, !LEAN, PT, !QG>(P, eb, pl, L);
  lower_prio<PRIO>();
  pencil_compute<NQ, ORDER, FMA, FACT_LIFT, NT, CV, LEAN, false, PT, QG, PRIO>(
      P, Drow, Lift, lfac, blk, pl, e, eb, act, L);
}

template <int NQ, int ORDER, bool FMA, bool FACT_LIFT, bool NT, bool CV>
__global__ void __launch_bounds__(64)
    pencil_kernel(const StageParams P, const double *__restrict__ Drow,
                  const double *__restrict__ Lift, const double *__restrict__ lfac) {
  using S = PencilShape<NQ>;
  constexpr int NFP = S::NFP, EPW = S::EPW;
  extern __shared__ double smem[];
  const int lane = threadIdx.x;
  const int el = lane / NFP;
  const int pl = lane - el * NFP;  // line index a + NQ*b

  // XCD-aware element order: workgroup b runs on XCD b % 8.  Linear order:
  // XCD x walks the x-th eighth of the element groups in order.  Tiled order
  // (structured meshes): see tiled_element.
  const int xcd = blockIdx.x % kNumXcd;
  const long long g = blockIdx.x / kNumXcd;
  long long e, e_any;
  if (P.tnx) {
    if (g * EPW >= P.tseq) return;  // uniform over the workgroup
    // the mapping's divisions, evaluated here: a table load would head each
    // wavefront's dependency chain (measured 1-2 % slower at p=7)
    e = tiled_element(P, xcd, g * EPW + el);
    // any valid element of this workgroup, for the address math of idle lanes
    e_any = e;
    for (int t = 0; t < EPW && e_any < 0; ++t) e_any = tiled_element(P, xcd, g * EPW + t);
    if (e_any < 0) return;  // uniform: no lane of this workgroup has an element
  } else {
    const long long ngroups = (P.Ne + EPW - 1) / EPW;
    const long long per = (ngroups + kNumXcd - 1) / kNumXcd;
    const long long gl = (long long)xcd * per + g;
    if (gl >= ngroups) return;  // uniform over the workgroup
    e = gl * EPW + el;
    if (e >= P.Ne) e = -1;
    e_any = gl * EPW;
  }
  pencil_element<NQ, ORDER, FMA, FACT_LIFT, NT, CV>(P, Drow, Lift, lfac,
                                                     smem + el * S::PER_ELEM, pl, e, e_any);
}

// Multi-wavefront pencil stage for Nq >= 6: Nq^2 lanes per element, epb
// elements per workgroup (a launch argument; the workgroup has
// pencil_group_threads(epb) threads, at most MAXTH), lean LDS layout.
// Element order as pencil_kernel, in groups of epb.
template <int NQ, int MAXTH, int ORDER, bool FMA, bool FACT_LIFT, bool NT, bool CV>
__global__ void __launch_bounds__(MAXTH)
    pencil_group_kernel(const StageParams P, const double *__restrict__ Drow,
                        const double *__restrict__ Lift, const double *__restrict__ lfac,
                        int epb) {
  constexpr int NFP = NQ * NQ;
  constexpr int PER = PencilShape<NQ>::lean_elem(CV);
  extern __shared__ double smem[];
  const int tid = threadIdx.x;
  const bool act = tid < epb * NFP;
  const int el = act ? tid / NFP : 0;
  const int pl = act ? tid - el * NFP : 0;
  const int xcd = blockIdx.x % kNumXcd;
  const long long g = blockIdx.x / kNumXcd;
  long long e, e_any;
  if (P.tnx) {
    if (g * epb >= P.tseq) return;  // uniform over the workgroup
    e = act ? tiled_lookup(P, xcd, g * epb + el) : -1;
    e_any = -1;
    for (int t = 0; t < epb && e_any < 0; ++t) e_any = tiled_lookup(P, xcd, g * epb + t);
    if (e_any < 0) return;  // uniform: no element in this workgroup
  } else {
    const long long ngroups = (P.Ne + epb - 1) / epb;
    const long long per = (ngroups + kNumXcd - 1) / kNumXcd;
    const long long gl = (long long)xcd * per + g;
    if (g >= per || gl >= ngroups) return;  // uniform over the workgroup
    e = gl * epb + el;
    if (!act || e >= P.Ne) e = -1;
    e_any = gl * epb;
  }
  pencil_element<NQ, ORDER, FMA, FACT_LIFT, NT, CV, true, const StageParams,
                 PencilShape<NQ>::lean_qg(CV)>(P, Drow, Lift, lfac, smem + el * PER, pl, e,
                                               e_any, act);
}

}  // namespace dgk
