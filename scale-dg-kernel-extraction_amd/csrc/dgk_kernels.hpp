// HIP device kernels of the DG advection tendency and SSP-RK3 stage for
// gfx950.  Compiled with -ffp-contract=off: every multiply-add is written out
// either as an explicit fma() (FMA variants) or as a rounded product followed
// by a rounded sum in the reference's order (exact variants), so that an
// exact variant reproduces the reference's FP64 results bit for bit.
//
// Element blocks are the reference's: q(Np,NeA) column-major, node
// n = i + j*Nq + k*Nq^2 inside element e at e*Np + n (mod_mesh.f90:267).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dgk {

constexpr int kGeneral = 1;  // mod_dg_optr_kernel.f90:39
constexpr int kOpt1 = 2;     // mod_dg_optr_kernel.f90:42
// Left-to-right sums at every p (no split): the OPT1 module's dirX/dirY/dirZ
// and gradlike kernels (mod_dg_optr_kernel_opt1.F90.erb:51-129, 166-219).
constexpr int kSequential = 3;
// Mixed precision (pencil kernels only): the directional contractions in
// FP32 with every operator entry split into hi + lo FP32 parts, applied to
// each line's remainder after its cubic part (taken out and differentiated
// in FP64); the face-flux differences, the lift, the residual sum and the RK
// update stay FP64.
constexpr int kMixed = 4;
constexpr int kNumXcd = 8;   // MI355X: 8 XCDs, workgroups dispatched round-robin

// Stage modes of stage_kernel.
constexpr int kModeTend = 0;    // write dqdt only
constexpr int kModeStage1 = 1;  // q <- rk_a q + rk_b (q + dt dqdt); the input becomes q0
constexpr int kModeStageN = 2;  // q <- rk_a q0 + rk_b (q + dt dqdt), q0 read

struct StageParams {
  const double *qin;   // (Np,NeA) current state, halo valid in table mode
  double *qout;        // (Np,NeA) next state (owned part written)
  const double *q0;    // (Np,NeA) state at the start of the step (stages 2, 3)
  const double *u, *v, *w;
  double *dqdt;        // (Np,NeA) tendency (kModeTend)
  const double *D;     // (Nq,Nq) D1D column-major
  const double *Lift;  // (Nq,Nq,Nq,6)
  // compressed connectivity: 0-based neighbour element per (face, element)
  const int32_t *nbr;  // (6,Ne)
  // per-node tables (reference layout)
  const int32_t *vmapM, *vmapP;  // (NfpTot,Ne) 1-based
  const double *nrm;             // (NfpTot,Ne,3)
  const double *fscale;          // (NfpTot,Ne)
  const double *escale;          // (Np,Ne,3)
  double fs[6];                  // compressed: Fscale per face
  double es[3];                  // compressed: Escale per axis
  double cvel[3];                // constant velocity (pencil kernel, CV)
  long long Ne;
  double a, b, dt;
  int mode;
  // Tiled element order (pencil kernel) for a periodic NeX x NeY x NeZ mesh
  // in the reference's numbering: tnx = 0 selects the plain linear order.
  int tnx, tny, tnz;  // mesh extents
  int ttx, tty;       // tile extents in x and y
  long long tseq;     // sequence positions per XCD
  // tiled_element(x, s) for every XCD x and position s < tseq (row x of
  // kNumXcd rows), built on the host: the multi-wavefront kernel looks its
  // elements up instead of evaluating the divisions of the mapping (3-6 %
  // faster at p=5 and 11)
  const int32_t *tile_order;
};

// Element at sequence position s of XCD x in the tiled order, or -1.
// XCD x owns the z-slab k in [x*NeZ/8, (x+1)*NeZ/8) and walks it tile by
// tile (ttx x tty elements in x,y), each tile plane by plane, so that the
// x-, y- and z-neighbours of an element are processed within one tile plane
// (ttx*tty elements) of it by the same XCD, whose L2 still holds their data.
__host__ __device__ __forceinline__ long long tiled_element(const StageParams &P, int x,
                                                            long long s) {
  const int k0 = (int)((long long)x * P.tnz / kNumXcd);
  const int k1 = (int)((long long)(x + 1) * P.tnz / kNumXcd);
  const int nk = k1 - k0;
  const int plane = P.ttx * P.tty;
  const long long per_tile = (long long)plane * nk;
  if (nk <= 0 || s >= P.tseq) return -1;
  const long long t = s / per_tile;
  const long long r = s - t * per_tile;
  const int k = k0 + (int)(r / plane);
  const int rr = (int)(r - (long long)(k - k0) * plane);
  const int jl = rr / P.ttx, il = rr - jl * P.ttx;
  const int ntx = (P.tnx + P.ttx - 1) / P.ttx;
  const long long tj = t / ntx, ti = t - tj * ntx;
  const long long i = ti * P.ttx + il, j = tj * P.tty + jl;
  if (i >= P.tnx || j >= P.tny) return -1;
  return i + (long long)P.tnx * (j + (long long)P.tny * k);
}

// The tiled order's element at position s of XCD x, from the host-built table.
__device__ __forceinline__ long long tiled_lookup(const StageParams &P, int x, long long s) {
  return s < P.tseq ? P.tile_order[(long long)x * P.tseq + s] : -1;
}

// q0 is read once per stage and by no neighbour: a non-temporal load keeps
// it from displacing lines that neighbouring elements read again from L2.
__device__ __forceinline__ double ld_stream(const double *p) {
  return __builtin_nontemporal_load(p);
}

// ---------------------------------------------------------------------------
// Arithmetic helpers.  madd(x,y,s) = s + x*y; FMA variants fuse it.
template <bool FMA>
__device__ __forceinline__ double madd(double x, double y, double s) {
  if constexpr (FMA) return __builtin_fma(x, y, s);
  else return s + x * y;
}

// Sum over l = 0..NQ-1 of a(l)*b(l) in the reference's order:
//   GENERAL: s = 0; s = s + a(l)*b(l)            (mod_dg_optr_kernel.f90:216-256)
//   OPT1   : left-to-right for p <= 7; for p > 7 three partial sums over
//            [0,r), [r,2r), [2r,NQ) with r = round(NQ/3), then t1 + t2 + t3
//            (mod_dg_optr_kernel_opt1.F90.erb:42-44, 249-342).
//   SEQUENTIAL: left-to-right at every p.
template <int NQ, int ORDER, bool FMA, class FA, class FB>
__device__ __forceinline__ double contract(FA a, FB b) {
  if constexpr (ORDER == kGeneral) {
    double s = 0.0;
#pragma unroll
    for (int l = 0; l < NQ; ++l) s = madd<FMA>(a(l), b(l), s);
    return s;
  } else if constexpr (NQ <= 8 || ORDER == kSequential) {
    double s = a(0) * b(0);
#pragma unroll
    for (int l = 1; l < NQ; ++l) s = madd<FMA>(a(l), b(l), s);
    return s;
  } else {
    constexpr int r = (3 * NQ + 4) / 9;  // round(NQ/3) for NQ = 9..16
    static_assert(r == 3 || r == 4 || r == 5, "split point");
    double t1 = a(0) * b(0);
#pragma unroll
    for (int l = 1; l < r; ++l) t1 = madd<FMA>(a(l), b(l), t1);
    double t2 = a(r) * b(r);
#pragma unroll
    for (int l = r + 1; l < 2 * r; ++l) t2 = madd<FMA>(a(l), b(l), t2);
    double t3 = a(2 * r) * b(2 * r);
#pragma unroll
    for (int l = 2 * r + 1; l < NQ; ++l) t3 = madd<FMA>(a(l), b(l), t3);
    return t1 + t2 + t3;
  }
}

// Face f (0-based: -y,+x,+y,-x,-z,+z) local node at tangential (a,b), first
// tangential index fastest (Fmask, mod_mesh.f90:278-324).
template <int NQ>
__device__ __forceinline__ int face_node(int f, int a, int b) {
  constexpr int NFP = NQ * NQ;
  switch (f) {
  case 0: return a + b * NFP;
  case 1: return (NQ - 1) + a * NQ + b * NFP;
  case 2: return a + (NQ - 1) * NQ + b * NFP;
  case 3: return a * NQ + b * NFP;
  case 4: return a + b * NQ;
  default: return a + b * NQ + (NQ - 1) * NFP;
  }
}

__device__ __forceinline__ int face_opp(int f) { return f ^ ((f < 4) ? 2 : 1); }
__device__ __forceinline__ int face_axis(int f) { return (f < 4) ? ((f & 1) ? 0 : 1) : 2; }
__device__ __forceinline__ bool face_plus(int f) { return f == 1 || f == 2 || f == 5; }

// Upwind flux on one face node, cal_elembnd_flux (mod_advect3d_eq.f90:117-139):
//   VelM/P = (u,v,w)(iM/iP) . n ; alpha = |VelP+VelM|/2
//   flux = 0.5*Fscale*(qP*VelP - qM*VelM - alpha*(qP-qM))
// With an axis-aligned unit normal (0 components exact zeros), VelM equals
// +-vel_axis(iM) exactly, which is what the compressed form evaluates.
__device__ __forceinline__ double upwind_flux(double qM, double qP, double VelM,
                                              double VelP, double Fs) {
  const double alpha = 0.5 * fabs(VelP + VelM);
  return 0.5 * Fs * (qP * VelP - qM * VelM - alpha * (qP - qM));
}

template <int NQ, bool STRUCT>
__device__ __forceinline__ double face_flux(const StageParams &P, long long e, int tt) {
  constexpr int NP = NQ * NQ * NQ, NFP = NQ * NQ, NFT = 6 * NFP;
  if constexpr (STRUCT) {
    const int f = tt / NFP, fp = tt - f * NFP;
    const int a = fp % NQ, b = fp / NQ;
    const long long iM = e * NP + face_node<NQ>(f, a, b);
    const long long eP = P.nbr[e * 6 + f];
    const long long iP = eP * NP + face_node<NQ>(face_opp(f), a, b);
    const int ax = face_axis(f);
    const double *vel = ax == 0 ? P.u : (ax == 1 ? P.v : P.w);
    const double vm = vel[iM], vp = vel[iP];
    const bool plus = face_plus(f);
    const double VelM = plus ? vm : -vm, VelP = plus ? vp : -vp;
    return upwind_flux(P.qin[iM], P.qin[iP], VelM, VelP, P.fs[f]);
  } else {
    const long long o = e * NFT + tt;
    const long long NfNe = (long long)NFT * P.Ne;
    const long long iM = P.vmapM[o] - 1, iP = P.vmapP[o] - 1;
    const double n1 = P.nrm[o], n2 = P.nrm[o + NfNe], n3 = P.nrm[o + 2 * NfNe];
    const double VelM = P.u[iM] * n1 + P.v[iM] * n2 + P.w[iM] * n3;
    const double VelP = P.u[iP] * n1 + P.v[iP] * n2 + P.w[iP] * n3;
    return upwind_flux(P.qin[iM], P.qin[iP], VelM, VelP, P.fscale[o]);
  }
}

// Launch shape of the stage kernel for a given Nq: EPB elements per
// workgroup, NT threads, NPT nodes per thread.
// At most 512 threads (8 wavefronts), so that the compiler may give each
// thread up to 256 VGPRs: with 1024-thread workgroups Nq >= 11 spilled.
template <int NQ>
struct StageShape {
  static constexpr int NP = NQ * NQ * NQ;
  static constexpr int EPB = NP >= 256 ? 1 : (256 / NP);
  static constexpr int NN = EPB * NP;
  static constexpr int NPT = (NN + 511) / 512;
  static constexpr int NT = ((NN + NPT - 1) / NPT + 63) / 64 * 64;
  static constexpr int NF = EPB * 6 * NQ * NQ;
  static constexpr size_t LDS = sizeof(double) * (3 * NN + NF + NQ * NQ);
};

// One fused SSP-RK3 stage (or a bare tendency) over all elements.
//
// Workgroups are persistent: blockIdx.x % 8 selects the XCD, and each XCD
// walks a contiguous range of element groups with its workgroups interleaved,
// so that x- and y-neighbours of an element were touched recently by the
// same XCD's L2.  Per group: (A) stage q*u, q*v, q*w in LDS and prefetch the
// next group's state into registers; (B) one upwind flux per face node into
// LDS; (C) per node the three sum-factorised sweeps, the lift and the RK
// combine, written straight to HBM.
template <int NQ, int ORDER, bool FMA, bool STRUCT>
__global__ void __launch_bounds__(StageShape<NQ>::NT)
    stage_kernel(const StageParams P) {
  using S = StageShape<NQ>;
  constexpr int NP = S::NP, NFP = NQ * NQ, NFT = 6 * NFP;
  constexpr int EPB = S::EPB, NT = S::NT, NN = S::NN, NPT = S::NPT, NF = S::NF;
  constexpr int NFPT = (NF + NT - 1) / NT;
  constexpr bool HOLD = NPT <= 2;  // Lift in registers + prefetch

  extern __shared__ double smem[];
  double *sfx = smem, *sfy = smem + NN, *sfz = smem + 2 * NN;
  double *sfb = smem + 3 * NN, *sD = smem + 3 * NN + NF;

  const int tid = threadIdx.x;
  const long long ngroups = (P.Ne + EPB - 1) / EPB;
  const int xcd = blockIdx.x % kNumXcd, slot = blockIdx.x / kNumXcd;
  const int nslot = gridDim.x / kNumXcd;
  const long long per = (ngroups + kNumXcd - 1) / kNumXcd;
  const long long gbeg = xcd * per;
  const long long gend = gbeg + per < ngroups ? gbeg + per : ngroups;

  for (int t = tid; t < NQ * NQ; t += NT) sD[t] = P.D[t];

  double lreg[HOLD ? NPT * 6 : 1];
  if constexpr (HOLD) {
#pragma unroll
    for (int r = 0; r < NPT; ++r) {
      const int m = tid + r * NT;
      const int n = (m < NN ? m : 0) % NP;
#pragma unroll
      for (int f = 0; f < 6; ++f) lreg[r * 6 + f] = P.Lift[n + f * NP];
    }
  }

  const bool need_q0 = P.mode == kModeStageN;
  double cq[NPT], cu[NPT], cv[NPT], cw[NPT], cq0[NPT];
  auto load = [&](long long g, double *xq, double *xu, double *xv, double *xw, double *xq0) {
#pragma unroll
    for (int r = 0; r < NPT; ++r) {
      const int m = tid + r * NT;
      const long long e = g * EPB + m / NP;
      if (m < NN && e < P.Ne) {
        const long long idx = g * NN + m;
        xq[r] = P.qin[idx];
        xu[r] = P.u[idx];
        xv[r] = P.v[idx];
        xw[r] = P.w[idx];
        xq0[r] = need_q0 ? ld_stream(P.q0 + idx) : 0.0;
      }
    }
  };

  long long g = gbeg + slot;
  if (HOLD && g < gend) load(g, cq, cu, cv, cw, cq0);

  for (; g < gend; g += nslot) {
    if (!HOLD) load(g, cq, cu, cv, cw, cq0);
    __syncthreads();  // LDS of the previous group fully consumed
    // (A) volume fluxes q*u, q*v, q*w (mod_advect3d_eq.f90:181-183)
#pragma unroll
    for (int r = 0; r < NPT; ++r) {
      const int m = tid + r * NT;
      if (m < NN && g * EPB + m / NP < P.Ne) {
        sfx[m] = cq[r] * cu[r];
        sfy[m] = cq[r] * cv[r];
        sfz[m] = cq[r] * cw[r];
      }
    }
    double nq[NPT], nu[NPT], nv[NPT], nw[NPT], nq0[NPT];
    if constexpr (HOLD) {
      if (g + nslot < gend) load(g + nslot, nq, nu, nv, nw, nq0);
    }
    // (B) element-boundary flux (mod_advect3d_eq.f90:116-140)
#pragma unroll
    for (int r = 0; r < NFPT; ++r) {
      const int t = tid + r * NT;
      if (t < NF) {
        const int el = t / NFT, tt = t - el * NFT;
        const long long e = g * EPB + el;
        if (e < P.Ne) sfb[t] = face_flux<NQ, STRUCT>(P, e, tt);
      }
    }
    __syncthreads();
    // (C) sweeps, lift, combine, RK (mod_advect3d_eq.f90:185-198, main.f90:69-73)
    // With many nodes per thread (Nq >= 14) the loop stays rolled and q, q0
    // are read again (from L2: this group loaded them in phase A), which
    // keeps the register count below the spill limit.
    auto node = [&](int r, double q, double q0v) {
      const int m = tid + r * NT;
      const long long e = g * EPB + m / NP;
      if (m < NN && e < P.Ne) {
        const int el = m / NP, n = m - el * NP;
        const int i = n % NQ, j = (n / NQ) % NQ, k = n / NFP;
        const double *bx = sfx + el * NP, *by = sfy + el * NP, *bz = sfz + el * NP;
        const double dx = contract<NQ, ORDER, FMA>(
            [&](int l) { return sD[i + l * NQ]; },
            [&](int l) { return bx[l + j * NQ + k * NFP]; });
        const double dy = contract<NQ, ORDER, FMA>(
            [&](int l) { return by[i + l * NQ + k * NFP]; },
            [&](int l) { return sD[j + l * NQ]; });
        const double dz = contract<NQ, ORDER, FMA>(
            [&](int l) { return bz[i + j * NQ + l * NFP]; },
            [&](int l) { return sD[k + l * NQ]; });
        const double *fb = sfb + el * NFT;
        double L[6];
#pragma unroll
        for (int f = 0; f < 6; ++f) L[f] = HOLD ? lreg[(HOLD ? r : 0) * 6 + f] : P.Lift[n + f * NP];
        // tensorprod_Lift_hexahedral (mod_dg_optr_kernel.f90:277-283)
        double lift = L[0] * fb[i + k * NQ];
        lift = madd<FMA>(L[1], fb[NFP + j + k * NQ], lift);
        lift = madd<FMA>(L[2], fb[2 * NFP + i + k * NQ], lift);
        lift = madd<FMA>(L[3], fb[3 * NFP + j + k * NQ], lift);
        lift = madd<FMA>(L[4], fb[4 * NFP + i + j * NQ], lift);
        lift = madd<FMA>(L[5], fb[5 * NFP + i + j * NQ], lift);
        const long long idx = g * NN + m;
        double e1, e2, e3;
        if constexpr (STRUCT) {
          e1 = P.es[0]; e2 = P.es[1]; e3 = P.es[2];
        } else {
          const long long NpNe = (long long)NP * P.Ne;
          e1 = P.escale[idx]; e2 = P.escale[idx + NpNe]; e3 = P.escale[idx + 2 * NpNe];
        }
        double s = e1 * dx;
        s = madd<FMA>(e2, dy, s);
        s = madd<FMA>(e3, dz, s);
        const double dq = -(s + lift);
        if (P.mode == kModeTend) {
          P.dqdt[idx] = dq;
        } else {
          const double q0s = P.mode == kModeStage1 ? q : q0v;
          double qn;
          if constexpr (FMA) qn = __builtin_fma(P.a, q0s, P.b * __builtin_fma(P.dt, dq, q));
          else qn = P.a * q0s + P.b * (q + P.dt * dq);
          P.qout[idx] = qn;
        }
      }
    };
    if constexpr (NPT <= 4) {
#pragma unroll
      for (int r = 0; r < NPT; ++r) node(r, cq[r], cq0[r]);
    } else {
#pragma unroll 1
      for (int r = 0; r < NPT; ++r) {
        const int m = tid + r * NT;
        const long long idx = g * NN + m;
        const bool ok = m < NN && g * EPB + m / NP < P.Ne;
        node(r, ok ? P.qin[idx] : 0.0, ok && need_q0 ? ld_stream(P.q0 + idx) : 0.0);
      }
    }
    if constexpr (HOLD) {
#pragma unroll
      for (int r = 0; r < NPT; ++r) {
        cq[r] = nq[r]; cu[r] = nu[r]; cv[r] = nv[r]; cw[r] = nw[r]; cq0[r] = nq0[r];
      }
    }
  }
}

// Batched tensorprod_divlike_dirXYZ: one element per workgroup.
template <int NQ, int ORDER>
__global__ void __launch_bounds__(256)
    divlike_batched_kernel(const double *Mat, const double *Mat_tr, const double *ix,
                           const double *iy, const double *iz, double *ox, double *oy,
                           double *oz) {
  constexpr int NP = NQ * NQ * NQ, NFP = NQ * NQ;
  extern __shared__ double smem[];
  double *sx = smem, *sy = smem + NP, *sz = smem + 2 * NP, *sM = smem + 3 * NP,
         *sT = smem + 3 * NP + NQ * NQ;
  const long long base = (long long)blockIdx.x * NP;
  for (int t = threadIdx.x; t < NP; t += blockDim.x) {
    sx[t] = ix[base + t];
    sy[t] = iy[base + t];
    sz[t] = iz[base + t];
  }
  for (int t = threadIdx.x; t < NQ * NQ; t += blockDim.x) {
    sM[t] = Mat[t];
    sT[t] = Mat_tr[t];
  }
  __syncthreads();
  for (int n = threadIdx.x; n < NP; n += blockDim.x) {
    const int i = n % NQ, j = (n / NQ) % NQ, k = n / NFP;
    ox[base + n] = contract<NQ, ORDER, false>([&](int l) { return sM[i + l * NQ]; },
                                              [&](int l) { return sx[l + j * NQ + k * NFP]; });
    oy[base + n] = contract<NQ, ORDER, false>([&](int l) { return sy[i + l * NQ + k * NFP]; },
                                              [&](int l) { return sT[l + j * NQ]; });
    oz[base + n] = contract<NQ, ORDER, false>([&](int l) { return sz[i + j * NQ + l * NFP]; },
                                              [&](int l) { return sT[l + k * NQ]; });
  }
}

// Batched tensorprod_Lift_hexahedral (mod_dg_optr_kernel.f90:262-288):
// persistent workgroups of NT threads.  Up to 256 nodes an element (Nq <= 6)
// a workgroup takes EPB = 256/Np elements at a time, one node per thread;
// above that one element at a time, NPT nodes per thread.  A thread's nodes
// never change, so their six Lift coefficients stay in registers and the
// table is read once per workgroup instead of once per element; the
// elements' face fluxes are staged in LDS.  Same products and order of
// additions as the reference.
template <int NQ>
struct LiftShape {
  static constexpr int NP = NQ * NQ * NQ, NFT = 6 * NQ * NQ;
  static constexpr int NT = NP <= 256 ? 256 : (NP <= 1024 ? (NP + 63) / 64 * 64 : 1024);
  static constexpr int EPB = NP <= 256 ? 256 / NP : 1;
  static constexpr int NPT = NP <= 256 ? 1 : (NP + NT - 1) / NT;
  static constexpr size_t LDS = sizeof(double) * EPB * NFT;
};

template <int NQ>
__global__ void __launch_bounds__(LiftShape<NQ>::NT)
    lift_batched_kernel(const double *Lift, const double *in, double *out, long long nelem) {
  using S = LiftShape<NQ>;
  constexpr int NP = S::NP, NFP = NQ * NQ, NFT = S::NFT, EPB = S::EPB, NPT = S::NPT;
  extern __shared__ double smem[];
  const int t = threadIdx.x;
  // this thread's element slot and nodes
  const int es = NP <= 256 ? t / NP : 0;
  int nn[NPT];
  double Lc[NPT][6];
#pragma unroll
  for (int r = 0; r < NPT; ++r) {
    const int n = NP <= 256 ? t % NP : t + r * S::NT;
    const bool ok = (NP <= 256 ? es < EPB : n < NP);
    nn[r] = ok ? n : -1;
#pragma unroll
    for (int f = 0; f < 6; ++f) Lc[r][f] = ok ? Lift[n + f * NP] : 0.0;
  }
  for (long long e0 = (long long)blockIdx.x * EPB; e0 < nelem; e0 += (long long)gridDim.x * EPB) {
    const long long nf = (nelem - e0 < EPB ? nelem - e0 : EPB) * NFT;
    for (int u = t; u < nf; u += S::NT) smem[u] = in[e0 * NFT + u];
    __syncthreads();
    const long long e = e0 + es;
    if (e < nelem) {
      const double *fb = smem + es * NFT;
#pragma unroll
      for (int r = 0; r < NPT; ++r) {
        const int n = nn[r];
        if (n < 0) continue;
        const int i = n % NQ, j = (n / NQ) % NQ, k = n / NFP;
        double s = Lc[r][0] * fb[i + k * NQ];
        s = s + Lc[r][1] * fb[NFP + j + k * NQ];
        s = s + Lc[r][2] * fb[2 * NFP + i + k * NQ];
        s = s + Lc[r][3] * fb[3 * NFP + j + k * NQ];
        s = s + Lc[r][4] * fb[4 * NFP + i + j * NQ];
        s = s + Lc[r][5] * fb[5 * NFP + i + j * NQ];
        out[e * NP + n] = s;
      }
    }
    __syncthreads();
  }
}

}  // namespace dgk
