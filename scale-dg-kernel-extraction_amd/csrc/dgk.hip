// C-ABI of the device-resident DG advection solver (include/dgk.h): device
// state, table verification/compression, launch dispatch over Nq and the
// summation order, halo gathers and the hipGraph-captured time step.
#include "dgk_internal.hpp"
#include "dgk_kernels.hpp"
#include "dgk_pencil.hpp"
#include "dgk_launch.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <vector>

using namespace dgk;

namespace dgk {

// update_halo (mod_mesh.f90:216-218): f(Np*Ne + ib) = f(halo_src_map(ib)).
// f[0..n) = v (a uniform velocity component, halo slots included).
__global__ void __launch_bounds__(256) fill_kernel(double *f, long long n, double v) {
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x)
    f[t] = v;
}

__global__ void halo_gather_kernel(double *f, const int32_t *src, int n, long long base) {
  const long long ib = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (ib < n) f[base + ib] = f[src[ib] - 1];
}

// min/max that keep a NaN operand (fmin/fmax drop it): a state that has gone
// NaN shows in the per-step line instead of vanishing from it.  On numbers
// they are fmin/fmax.
__device__ __forceinline__ double nan_min(double a, double b) {
  return (a != a || b != b) ? a + b : fmin(a, b);
}
__device__ __forceinline__ double nan_max(double a, double b) {
  return (a != a || b != b) ? a + b : fmax(a, b);
}

// Two-pass (min,max) of the owned state: per-block partials, then one block.
__global__ void __launch_bounds__(256)
    minmax_partial_kernel(const double *q, long long n, double *part) {
  __shared__ double smn[256], smx[256];
  double mn = __builtin_inf(), mx = -__builtin_inf();
  // unrolled so that several loads of a thread are in flight at once; a NaN
  // value sticks (no comparison with a NaN holds)
#pragma unroll 4
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    const double x = q[t];
    mn = (x < mn || x != x) ? x : mn;
    mx = (x > mx || x != x) ? x : mx;
  }
  smn[threadIdx.x] = mn;
  smx[threadIdx.x] = mx;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      smn[threadIdx.x] = nan_min(smn[threadIdx.x], smn[threadIdx.x + s]);
      smx[threadIdx.x] = nan_max(smx[threadIdx.x], smx[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = smn[0];
    part[2 * blockIdx.x + 1] = smx[0];
  }
}

__global__ void __launch_bounds__(256)
    minmax_final_kernel(const double *part, int nb, double *out) {
  __shared__ double smn[256], smx[256];
  double mn = __builtin_inf(), mx = -__builtin_inf();
  for (int t = threadIdx.x; t < nb; t += blockDim.x) {
    mn = nan_min(mn, part[2 * t]);
    mx = nan_max(mx, part[2 * t + 1]);
  }
  smn[threadIdx.x] = mn;
  smx[threadIdx.x] = mx;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      smn[threadIdx.x] = nan_min(smn[threadIdx.x], smn[threadIdx.x + s]);
      smx[threadIdx.x] = nan_max(smx[threadIdx.x], smx[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = smn[0];
    out[1] = smx[0];
  }
}

// Diagnostics of the owned state q (and of q - r for a reference field r),
// per-block partials then one block:
//   0 min q   1 max q   2 sum J w q   3 sum J w |q|   4 sum q^2
//   5 sum (q-r)^2   6 sum r^2   7 max |q-r|
// w = w_i w_j w_k (LGL weights of node n), J = 1/(Es1 Es2 Es3) = the element
// volume / 8: the sums 2 and 3 are the mass and the |q|-mass of the field.
// Either es (compressed: per-axis Escale) or escale (Np,Ne,3) is used.
constexpr int kDiag = 8;

struct DiagArgs {
  const double *q, *r;  // r may be null
  const double *escale; // (Np,Ne,3) or null
  double w1[16];        // LGL weights (Nq <= 16), passed by value
  double es[3];
  long long n, NpNe;
  int Nq, Np;
};

__device__ __forceinline__ void diag_combine(double *a, const double *b) {
  a[0] = nan_min(a[0], b[0]);
  a[1] = nan_max(a[1], b[1]);
#pragma unroll
  for (int t = 2; t < 7; ++t) a[t] += b[t];
  a[7] = nan_max(a[7], b[7]);
}

__device__ __forceinline__ void diag_block_reduce(double (&v)[kDiag], double *out) {
  __shared__ double sm[kDiag][256];
#pragma unroll
  for (int t = 0; t < kDiag; ++t) sm[t][threadIdx.x] = v[t];
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      double a[kDiag], b[kDiag];
#pragma unroll
      for (int t = 0; t < kDiag; ++t) {
        a[t] = sm[t][threadIdx.x];
        b[t] = sm[t][threadIdx.x + h];
      }
      diag_combine(a, b);
#pragma unroll
      for (int t = 0; t < kDiag; ++t) sm[t][threadIdx.x] = a[t];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0)
#pragma unroll
    for (int t = 0; t < kDiag; ++t) out[t] = sm[t][0];
}

__device__ __forceinline__ void diag_init(double (&v)[kDiag]) {
  v[0] = __builtin_inf();
  v[1] = -__builtin_inf();
#pragma unroll
  for (int t = 2; t < kDiag; ++t) v[t] = 0.0;
}

// Grid-stride over the owned nodes; a thread's node index within its element
// advances by the stride modulo Np, and the node's weight w_i w_j w_k comes
// from an LDS table of the Np weights (dynamic LDS, Np doubles), so the loop
// does no integer division.  REF: A.r is a reference field (sums 5-7); ESC:
// per-node Escale (A.escale) instead of the uniform A.es.
template <bool REF, bool ESC>
__global__ void __launch_bounds__(256) diag_partial_kernel(const DiagArgs A, double *part) {
  extern __shared__ double wn[];
  for (int n = threadIdx.x; n < A.Np; n += blockDim.x) {
    const int i = n % A.Nq, j = (n / A.Nq) % A.Nq, k = n / (A.Nq * A.Nq);
    wn[n] = A.w1[i] * A.w1[j] * A.w1[k];
  }
  __syncthreads();
  double v[kDiag];
  diag_init(v);
  const double J0 = 1.0 / (A.es[0] * A.es[1] * A.es[2]);
  const long long stride = (long long)gridDim.x * blockDim.x;
  const int dn = (int)(stride % A.Np);
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  int nd = (int)(t % A.Np);
  auto add = [&](double x, double r, double e0, double e1, double e2) {
    const double J = ESC ? 1.0 / (e0 * e1 * e2) : J0;
    const double wj = J * wn[nd];
    nd += dn;
    if (nd >= A.Np) nd -= A.Np;
    v[0] = nan_min(v[0], x);
    v[1] = nan_max(v[1], x);
    v[2] += wj * x;
    v[3] += wj * fabs(x);
    v[4] += x * x;
    if constexpr (REF) {
      const double d = x - r;
      v[5] += d * d;
      v[6] += r * r;
      v[7] = nan_max(v[7], fabs(d));
    }
  };
  // The values of kU strides are requested before any is reduced (left to
  // the compiler, each load waited before the next was issued, the LDS
  // weight lookup sitting between them); the reduction order is the
  // thread's stride order either way.
  constexpr int kU = 4;
  for (; t + (kU - 1) * stride < A.n; t += kU * stride) {
    double x[kU], r[kU], e[3][kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const long long tu = t + u * stride;
      x[u] = A.q[tu];
      r[u] = REF ? A.r[tu] : 0.0;
#pragma unroll
      for (int d = 0; d < 3; ++d) e[d][u] = ESC ? A.escale[tu + d * A.NpNe] : 1.0;
    }
    __builtin_amdgcn_sched_barrier(0);  // every load above is issued first
#pragma unroll
    for (int u = 0; u < kU; ++u) add(x[u], r[u], e[0][u], e[1][u], e[2][u]);
  }
  for (; t < A.n; t += stride)
    add(A.q[t], REF ? A.r[t] : 0.0, ESC ? A.escale[t] : 1.0, ESC ? A.escale[t + A.NpNe] : 1.0,
        ESC ? A.escale[t + 2 * A.NpNe] : 1.0);
  diag_block_reduce(v, part + kDiag * blockIdx.x);
}

__global__ void __launch_bounds__(256) diag_final_kernel(const double *part, int nb, double *out) {
  double v[kDiag];
  diag_init(v);
  for (int b = threadIdx.x; b < nb; b += blockDim.x) diag_combine(v, part + kDiag * b);
  diag_block_reduce(v, out);
}

}  // namespace dgk

#define HIPOK(expr)                                                            \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess)                                                      \
      return fail(DGK_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

namespace dgk {

// The two reduction kernels of the diagnostics on stream st: per-block
// partials into part (kDiag doubles per block, at most max_blocks blocks;
// every block writes its own slot before the final kernel reads it, so part
// needs no clearing), the eight results into res (device).
int diagnostics_launch(const double *q, const double *ref, long long n, int Nq,
                       const double es[3], const double *escale, const double *w1D,
                       double *part, int max_blocks, hipStream_t st, double *res) {
  if (Nq < 2 || Nq > 16) return fail(DGK_ERR_ARG, "Nq %d outside 2..16", Nq);
  DiagArgs A{};
  A.q = q;
  A.r = ref;
  for (int i = 0; i < Nq; ++i) A.w1[i] = w1D[i];
  A.escale = escale;
  for (int d = 0; d < 3; ++d) A.es[d] = es[d];
  A.n = n;
  A.NpNe = n;
  A.Nq = Nq;
  A.Np = Nq * Nq * Nq;
  const int nb = (int)std::min<long long>(max_blocks, (n + 255) / 256);
  if (nb < 1 || nb > max_blocks)
    return fail(DGK_ERR_ARG, "diagnostics: %d partial blocks for %lld values outside 1..%d", nb,
                n, max_blocks);
  auto kern = ref ? (escale ? diag_partial_kernel<true, true> : diag_partial_kernel<true, false>)
                 : (escale ? diag_partial_kernel<false, true> : diag_partial_kernel<false, false>);
  hipLaunchKernelGGL(kern, dim3(nb), dim3(256), sizeof(double) * A.Np, st, A, part);
  HIPOK(hipGetLastError());
  hipLaunchKernelGGL(diag_final_kernel, dim3(1), dim3(256), 0, st, part, nb, res);
  HIPOK(hipGetLastError());
  return DGK_OK;
}

int device_diagnostics(const double *q, const double *ref, long long n, int Nq,
                       const double es[3], const double *escale, const double *w1D,
                       double *scratch, void *stream, double *out) {
  constexpr int kMaxBlocks = 1024;
  static_assert(kDiag * kMaxBlocks + kDiag == kDiagScratch, "scratch layout");
  hipStream_t st = (hipStream_t)stream;
  double *part = scratch, *res = scratch + kDiag * kMaxBlocks;
  int rc = diagnostics_launch(q, ref, n, Nq, es, escale, w1D, part, kMaxBlocks, st, res);
  if (rc) return rc;
  HIPOK(hipMemcpyAsync(out, res, kDiag * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPOK(hipStreamSynchronize(st));
  return DGK_OK;
}

// Scratch of dgk_diagnostics_noalloc: the per-block partials of the reduction
// (kDiag x kNoAllocBlocks doubles, 64 KiB).  Each GPU the library runs on
// holds its own copy until the process ends.  In host code `scratch` only
// names the symbol: the device address of a GPU's copy comes from
// hipGetSymbolAddress with that GPU current (noalloc_scratch).
constexpr int kNoAllocBlocks = 1024;
static __device__ __align__(256) double scratch[kDiag * kNoAllocBlocks];

constexpr int kMaxDevices = 64;

// Device address of `scratch` on GPU dev, looked up once per device (with
// dev current) and kept.
int noalloc_scratch(int dev, double **ptr) {
  static std::mutex mu;
  static double *addr[kMaxDevices] = {};
  if (dev < 0 || dev >= kMaxDevices)
    return fail(DGK_ERR_ARG, "device %d outside 0..%d", dev, kMaxDevices - 1);
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[dev]) {
    int cur = -1;
    HIPOK(hipGetDevice(&cur));
    if (cur != dev) HIPOK(hipSetDevice(dev));
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) HIPOK(hipSetDevice(cur));
    if (e != hipSuccess || !p)
      return fail(DGK_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
    addr[dev] = (double *)p;
  }
  *ptr = addr[dev];
  return DGK_OK;
}

}  // namespace dgk

namespace {

using StageFn = void (*)(const StageParams);

template <int NQ, int ORDER, bool FMA, bool STRUCT>
void launch_node(const StageKernel &k, const StageParams &P, const HostOps &, const double *,
                 hipStream_t st) {
  hipLaunchKernelGGL((stage_kernel<NQ, ORDER, FMA, STRUCT>), dim3((unsigned)k.grid),
                     dim3(k.nt), k.lds, st, P);
}

template <int NQ, int ORDER, bool FMA, bool STRUCT>
StageKernel make_stage() {
  StageKernel k;
  k.kernel = (const void *)stage_kernel<NQ, ORDER, FMA, STRUCT>;
  k.launch = launch_node<NQ, ORDER, FMA, STRUCT>;
  k.nt = StageShape<NQ>::NT;
  k.lds = StageShape<NQ>::LDS;
  k.epb = StageShape<NQ>::EPB;
  return k;
}

template <int NQ, int ORDER, bool FMA, bool FACT, bool NT, bool CV>
void launch_pencil(const StageKernel &k, const StageParams &P, const HostOps &H,
                   const double *Lift, hipStream_t st) {
  (void)Lift;
  const double *dr = ORDER == kMixed ? H.Drow32 : H.Drow;
  const double *lf = H.lfac;
  hipLaunchKernelGGL((pencil_kernel<NQ, ORDER, FMA, FACT, NT, CV>), dim3((unsigned)k.grid),
                     dim3(64), k.lds, st, P, dr, H.LiftP, lf);
}

template <int NQ, int ORDER, bool FMA, bool FACT, bool NT, bool CV>
StageKernel make_pencil() {
  StageKernel k;
  k.kernel = (const void *)pencil_kernel<NQ, ORDER, FMA, FACT, NT, CV>;
  k.launch = launch_pencil<NQ, ORDER, FMA, FACT, NT, CV>;
  k.nt = 64;
  k.lds = PencilShape<NQ>::LDS;
  k.epb = PencilShape<NQ>::EPW;
  k.pencil = true;
  return k;
}

// Elements per workgroup of the group pencil kernel, narrow / wide (the
// workgroup is the lanes rounded up to whole wavefronts): Nq 2, 4, 8 (wide
// only) one wavefront, as pencil_kernel but in the lean layout; Nq 3 28 / 64;
// Nq 5 10 / 23; Nq 6 7 / 16; Nq 7 5 / 11; Nq 9 2 / 3; Nq 10 2 / 3; Nq 11 1
// / 2; Nq 12 1 / 2; Nq 13..16 one shape, 1 element.
int group_epb(int Nq, bool wide) {
  switch (Nq) {
  case 3: return wide ? 64 : 28;
  case 5: return wide ? 23 : 10;
  case 6: return wide ? 16 : 7;
  case 7: return wide ? 11 : 5;
  case 9: return wide ? 3 : 2;
  case 10: return wide ? 3 : 2;
  case 11: return wide ? 2 : 1;
  case 12: return wide ? 2 : 1;
  case 2: case 4: case 8: return 64 / (Nq * Nq);
  default: return 1;
  }
}

// Runtime choices to template arguments, one level at a time.
template <int NQ, int ORDER, bool FMA, bool FACT, bool NT>
StageKernel pick_pencil_cv(bool cv) {
  return cv ? make_pencil<NQ, ORDER, FMA, FACT, NT, true>()
            : make_pencil<NQ, ORDER, FMA, FACT, NT, false>();
}
template <int NQ, int ORDER, bool FMA, bool FACT>
StageKernel pick_pencil_nt(bool nt, bool cv) {
  return nt ? pick_pencil_cv<NQ, ORDER, FMA, FACT, true>(cv)
            : pick_pencil_cv<NQ, ORDER, FMA, FACT, false>(cv);
}
template <int NQ, int ORDER, bool FMA>
StageKernel pick_pencil_fact(bool fact, bool nt, bool cv) {
  return fact ? pick_pencil_nt<NQ, ORDER, FMA, true>(nt, cv)
              : pick_pencil_nt<NQ, ORDER, FMA, false>(nt, cv);
}
template <int NQ>
StageKernel pick_pencil_nq(int order, bool fma, bool fact, bool nt, bool cv) {
  if (order == kMixed) {  // FP32 split contractions: plain products, factored lift only
    if (fma || !fact) return StageKernel{};
    return pick_pencil_nt<NQ, kMixed, false, true>(nt, cv);
  }
  if (order == kGeneral)
    return fma ? pick_pencil_fact<NQ, kGeneral, true>(fact, nt, cv)
               : pick_pencil_fact<NQ, kGeneral, false>(fact, nt, cv);
  return fma ? pick_pencil_fact<NQ, kOpt1, true>(fact, nt, cv)
             : pick_pencil_fact<NQ, kOpt1, false>(fact, nt, cv);
}

bool pencil_nq(int Nq) { return Nq >= 2 && Nq <= 16; }

// Nq 2, 4, 8: pencil_kernel, or with `wide` the group kernel (one
// wavefront, lean layout); Nq 6, 12, 16: the group kernel.
StageKernel pick_pencil(int Nq, int order, bool fma, bool fact, bool nt, bool cv, bool wide) {
  const bool one_wave = Nq == 2 || Nq == 4 || Nq == 8;
  if (order == kMixed && one_wave && wide) return StageKernel{};
  if (!one_wave || wide)
    return pick_pencil_group(Nq, order, fma, fact, cv, group_epb(Nq, wide));
  switch (Nq) {
  case 2: return pick_pencil_nq<2>(order, fma, fact, nt, cv);
  case 4: return pick_pencil_nq<4>(order, fma, fact, nt, cv);
  case 8: return pick_pencil_nq<8>(order, fma, fact, nt, cv);
  default: return StageKernel{};
  }
}

template <int NQ>
StageKernel pick_stage_nq(int order, bool fma, bool structured) {
  const bool gen = order == kGeneral;
  if (structured) {
    if (gen) return fma ? make_stage<NQ, kGeneral, true, true>() : make_stage<NQ, kGeneral, false, true>();
    return fma ? make_stage<NQ, kOpt1, true, true>() : make_stage<NQ, kOpt1, false, true>();
  }
  if (gen) return fma ? make_stage<NQ, kGeneral, true, false>() : make_stage<NQ, kGeneral, false, false>();
  return fma ? make_stage<NQ, kOpt1, true, false>() : make_stage<NQ, kOpt1, false, false>();
}

StageKernel pick_stage(int Nq, int order, bool fma, bool structured) {
  switch (Nq) {
  case 2: return pick_stage_nq<2>(order, fma, structured);
  case 3: return pick_stage_nq<3>(order, fma, structured);
  case 4: return pick_stage_nq<4>(order, fma, structured);
  case 5: return pick_stage_nq<5>(order, fma, structured);
  case 6: return pick_stage_nq<6>(order, fma, structured);
  case 7: return pick_stage_nq<7>(order, fma, structured);
  case 8: return pick_stage_nq<8>(order, fma, structured);
  case 9: return pick_stage_nq<9>(order, fma, structured);
  case 10: return pick_stage_nq<10>(order, fma, structured);
  case 11: return pick_stage_nq<11>(order, fma, structured);
  case 12: return pick_stage_nq<12>(order, fma, structured);
  case 13: return pick_stage_nq<13>(order, fma, structured);
  case 14: return pick_stage_nq<14>(order, fma, structured);
  case 15: return pick_stage_nq<15>(order, fma, structured);
  case 16: return pick_stage_nq<16>(order, fma, structured);
  default: return StageKernel{};
  }
}

int order_of(int typeid_) {
  if (typeid_ == DGK_TYPE_GENERAL) return kGeneral;
  if (typeid_ == DGK_TYPE_OPT1 || typeid_ == DGK_TYPE_HIP) return kOpt1;
  return -1;
}

const double kRkA[3] = {0.0, 0.75, 1.0 / 3.0};  // mod_common.f90:29
const double kRkB[3] = {1.0, 0.25, 2.0 / 3.0};  // mod_common.f90:30

template <class T>
int dalloc(T **p, size_t n) {
  *p = nullptr;
  if (n == 0) return DGK_OK;
  hipError_t e = hipMalloc((void **)p, n * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(DGK_ERR_NOMEM, "hipMalloc of %zu bytes failed: %s", n * sizeof(T),
                hipGetErrorString(e));
  }
  return DGK_OK;
}

}  // namespace

struct dgk_solver {
  int Nq = 0, Np = 0, Nfp = 0, NfpTot = 0;
  long long Ne = 0, NeA = 0;
  int NhaloNode = 0;
  int order = kOpt1;
  bool fma = false;
  bool structured = false;
  int flags = 0;
  int variant = -1;  // pencil tuning variant; -1 = built-in default
  int NeX = 0, NeY = 0, NeZ = 0;  // structured extents in the reference numbering (0: unknown)
  double cvel[3] = {0, 0, 0};     // uniform velocity (DGK_FLAG_CONST_VELOCITY)
  HostOps ops;
  // State buffers: q and q0 rotate among three, so that no stage copies q
  // to q0 (stage 1 takes the current buffer as q0; every stage writes the
  // buffer that holds neither its input nor q0).
  double *q[3] = {nullptr, nullptr, nullptr};
  double *u = nullptr, *v = nullptr, *w = nullptr, *dqdt = nullptr;
  double *D = nullptr, *Lift = nullptr, *lfac = nullptr, *Drow = nullptr, *LiftP = nullptr;
  double *Drow32 = nullptr;  // FP32 split operator and line-fit tables (mixed precision)
  bool lift_packed = false;  // LiftP holds the delta-packed Lift (pack_lift_deltas)
  int32_t *nbr = nullptr, *vmapM = nullptr, *vmapP = nullptr, *halo = nullptr;
  double *nrm = nullptr, *fscale = nullptr, *escale = nullptr;
  double fs[6] = {0, 0, 0, 0, 0, 0}, es[3] = {0, 0, 0};
  double *part = nullptr, *mm = nullptr;
  double *diag = nullptr;  // diagnostics scratch: partials (8 x 1024), result (8)
  int32_t *tile_order = nullptr;  // tiled element order (StageParams::tile_order)
  int mm_cap = 0;
  int cur = 0;  // buffer of q
  int c0 = 2;   // buffer of q0 (never cur)
  hipStream_t stream = nullptr;
  int device = 0;  // the GPU current at creation (the stream's)
  StageKernel sk;
  // one captured step per (cur, c0) it starts from, and the (cur, c0) it ends in
  hipGraphExec_t graph[9] = {};
  int graph_next[9][2] = {};
  double graph_dt = 0.0;
  bool fields_set = false;
};

namespace {

// Built-in default tuning variant (overridden by dgk_set_variant), from
// sweeps of every variant at 16^3-64^3: the linear element order with
// non-temporal stores, except
// - the tiled order at Nq = 3 and 12, and at Nq = 8 with per-node velocity
//   (p=7, 64^3: 1230 against 1335 us; p=11: 2-4 % faster with the factored
//   lift and with uniform velocity), where the neighbours' face traces are
//   the larger share of the traffic;
// - plain stores at Nq = 2 (64-byte elements, 16 per wavefront) and at
//   Nq = 4 with uniform velocity (p=3, 64^3: 90.6 against 98.4 us).
// Linear against tiled elsewhere: p=1 6-15 %, p=3 6-19 %, p=4 3 %,
// p=5-15 0-4 % faster.
int default_pencil_variant(const dgk_solver *s) {
  const bool cv = (s->flags & DGK_FLAG_CONST_VELOCITY) != 0;
  switch (s->Nq) {
  case 2: return kPencilLinearOrder;
  case 4: return cv ? kPencilLinearOrder : kPencilNtStore | kPencilLinearOrder;
  case 3:
  case 12: return kPencilNtStore;
  case 8: return cv && s->order != kMixed ? kPencilNtStore | kPencilLinearOrder : kPencilNtStore;
  default: return kPencilNtStore | kPencilLinearOrder;
  }
}

int current_variant(const dgk_solver *s) {
  return s->variant >= 0 ? s->variant : default_pencil_variant(s);
}

// Tiled element order of the pencil kernel (tiled_element): on for a known
// structured numbering with at least one z-plane per XCD, unless the tuning
// variant asks for the linear order.
bool tiled_order(const dgk_solver *s) {
  return s->sk.pencil && s->NeX > 0 && s->NeZ >= kNumXcd &&
         !(current_variant(s) & kPencilLinearOrder);
}

// Tiles of up to 32x32 elements: x-neighbours' face traces (a 64-byte-
// strided gather that touches every line of the neighbour) are then read from
// L2 for all but 1/32 of the elements; at 64^3, p=7 tiles 16x8, 32x8, 32x16
// and 64x16 measured 1-8 % slower.  The extent is split into equal tiles
// (48 -> 2 x 24), so that few tile slots fall outside the mesh.
void tile_shape(const dgk_solver *s, int *tx, int *ty, long long *tseq) {
  auto even = [](int n) { return (n + (n + 31) / 32 - 1) / ((n + 31) / 32); };
  *tx = even(s->NeX);
  *ty = even(s->NeY);
  const long long ntx = (s->NeX + *tx - 1) / *tx, nty = (s->NeY + *ty - 1) / *ty;
  const long long nk = (s->NeZ + kNumXcd - 1) / kNumXcd;
  *tseq = ntx * nty * (long long)(*tx) * (*ty) * nk;
}

// Tuning variants of the solver's stage kernel: store flavour x element
// order x workgroup shape (Nq 13-16 have one shape).
int pencil_variants(const dgk_solver *s) {
  if (!s->sk.pencil) return 1;
  if (s->flags & DGK_FLAG_MFMA) return kPencilVariants;  // four or eight wavefronts
  return (s->Nq >= 13 || s->order == kMixed) ? kPencilWide : kPencilVariants;
}

template <class T>
int upload(T **dst, const T *src, size_t n, hipStream_t st);

int configure_launch(dgk_solver *s) {
  const bool use_pencil = s->structured && !(s->flags & DGK_FLAG_NODE_KERNEL) &&
                          pencil_nq(s->Nq) &&
                          (s->lift_packed || (s->flags & DGK_FLAG_FACTORED_LIFT) ||
                           s->order == kMixed);
  const int var = current_variant(s);
  if (s->order == kMixed && !use_pencil)
    return fail(DGK_ERR_ARG, "DGK_FLAG_MIXED needs the pencil kernel (compressed "
                             "connectivity, no DGK_FLAG_NODE_KERNEL) without DGK_FLAG_FMA");
  if ((s->flags & DGK_FLAG_CONST_VELOCITY) && !use_pencil && !s->lift_packed &&
      !(s->flags & DGK_FLAG_FACTORED_LIFT))
    return fail(DGK_ERR_ARG, "DGK_FLAG_CONST_VELOCITY needs the pencil kernel, whose exact lift "
                             "needs a Lift whose values along each face's normal index lie within "
                             "255 units in the last place of one another: add "
                             "DGK_FLAG_FACTORED_LIFT");
  if ((s->flags & DGK_FLAG_CONST_VELOCITY) && !use_pencil)
    return fail(DGK_ERR_ARG, "DGK_FLAG_CONST_VELOCITY needs the pencil kernel (compressed "
                             "connectivity, no DGK_FLAG_NODE_KERNEL)");
  if (s->flags & DGK_FLAG_MFMA) {
    if (!use_pencil || s->Nq != 16 || !(s->flags & DGK_FLAG_CONST_VELOCITY))
      return fail(DGK_ERR_ARG, "DGK_FLAG_MFMA needs PolyOrder 15, DGK_FLAG_CONST_VELOCITY and "
                               "compressed connectivity (no DGK_FLAG_NODE_KERNEL)");
    s->sk = pick_mfma(s->Nq, s->order == kMixed, (var & kPencilWide) != 0);
  } else if (use_pencil) {
    s->sk = pick_pencil(s->Nq, s->order, s->fma, (s->flags & DGK_FLAG_FACTORED_LIFT) != 0,
                        (var & kPencilNtStore) != 0, (s->flags & DGK_FLAG_CONST_VELOCITY) != 0,
                        (var & kPencilWide) != 0);
  } else {
    s->sk = pick_stage(s->Nq, s->order, s->fma, s->structured);
  }
  if (!s->sk.kernel && s->order == kMixed)
    return fail(DGK_ERR_ARG, "DGK_FLAG_MIXED needs the pencil kernel (compressed "
                             "connectivity, no DGK_FLAG_NODE_KERNEL) without DGK_FLAG_FMA");
  if (!s->sk.kernel) return fail(DGK_ERR_ORDER, "Nq %d not supported", s->Nq);
  if ((s->flags & DGK_FLAG_FACTORED_LIFT) && !use_pencil)
    return fail(DGK_ERR_ARG, "DGK_FLAG_FACTORED_LIFT needs the pencil kernel (compressed "
                             "connectivity, no DGK_FLAG_NODE_KERNEL)");
  HIPOK(hipFuncSetAttribute(s->sk.kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)s->sk.lds));
  const long long ngroups = (s->Ne + s->sk.epb - 1) / s->sk.epb;
  if (s->sk.pencil) {
    // one workgroup per element group, remapped per XCD inside the kernel
    if (tiled_order(s)) {
      int tx, ty;
      long long tseq;
      tile_shape(s, &tx, &ty, &tseq);
      s->sk.grid = (tseq + s->sk.epb - 1) / s->sk.epb * kNumXcd;
      if (!s->tile_order) {
        // the tiled order, evaluated once on the host (it depends on the
        // mesh extents only)
        StageParams T{};
        T.tnx = s->NeX;
        T.tny = s->NeY;
        T.tnz = s->NeZ;
        T.ttx = tx;
        T.tty = ty;
        T.tseq = tseq;
        std::vector<int32_t> tab((size_t)kNumXcd * tseq);
        for (int x = 0; x < kNumXcd; ++x)
          for (long long q = 0; q < tseq; ++q)
            tab[(size_t)x * tseq + q] = (int32_t)tiled_element(T, x, q);
        int rc = upload(&s->tile_order, tab.data(), tab.size(), s->stream);
        if (rc) return rc;
      }
    } else {
      s->sk.grid = (ngroups + kNumXcd - 1) / kNumXcd * kNumXcd;
    }
    return DGK_OK;
  }
  int dev = 0, ncu = 0, per_cu = 0;
  HIPOK(hipGetDevice(&dev));
  HIPOK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  HIPOK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, s->sk.kernel, s->sk.nt,
                                                     s->sk.lds));
  if (per_cu < 1) per_cu = 1;
  long long grid = (long long)per_cu * ncu;
  const long long need = (ngroups + kNumXcd - 1) / kNumXcd * kNumXcd;
  if (grid > need) grid = need;
  s->sk.grid = std::max<long long>(kNumXcd, grid / kNumXcd * kNumXcd);
  return DGK_OK;
}

int alloc_fields(dgk_solver *s) {
  const size_t nf = (size_t)s->Np * (size_t)s->NeA;
  int rc;
  if ((rc = dalloc(&s->q[0], nf)) || (rc = dalloc(&s->q[1], nf)) || (rc = dalloc(&s->q[2], nf)) ||
      (rc = dalloc(&s->u, nf)) || (rc = dalloc(&s->v, nf)) || (rc = dalloc(&s->w, nf)) ||
      (rc = dalloc(&s->dqdt, nf)))
    return rc;
  HIPOK(hipMemsetAsync(s->q[2], 0, nf * sizeof(double), s->stream));
  HIPOK(hipMemsetAsync(s->dqdt, 0, nf * sizeof(double), s->stream));
  HIPOK(hipMemsetAsync(s->q[1], 0, nf * sizeof(double), s->stream));
  if ((rc = dalloc(&s->part, 2 * 1024))) return rc;
  return DGK_OK;
}

int launch_halo(dgk_solver *s, double *f) {
  if (s->NhaloNode == 0) return DGK_OK;
  const int nb = (s->NhaloNode + 255) / 256;
  hipLaunchKernelGGL(halo_gather_kernel, dim3(nb), dim3(256), 0, s->stream, f, s->halo,
                     s->NhaloNode, (long long)s->Np * s->Ne);
  HIPOK(hipGetLastError());
  return DGK_OK;
}

StageParams make_params(dgk_solver *s, int mode, double a, double b, double dt) {
  StageParams P{};
  if (tiled_order(s)) {
    P.tnx = s->NeX;
    P.tny = s->NeY;
    P.tnz = s->NeZ;
    tile_shape(s, &P.ttx, &P.tty, &P.tseq);
    P.tile_order = s->tile_order;
  }
  P.qin = s->q[s->cur];
  P.qout = s->q[3 - s->cur - s->c0];
  P.q0 = s->q[mode == kModeStage1 ? s->cur : s->c0];
  P.u = s->u;
  P.v = s->v;
  P.w = s->w;
  P.dqdt = s->dqdt;
  P.D = s->D;
  P.Lift = s->Lift;
  P.nbr = s->nbr;
  P.vmapM = s->vmapM;
  P.vmapP = s->vmapP;
  P.nrm = s->nrm;
  P.fscale = s->fscale;
  P.escale = s->escale;
  for (int f = 0; f < 6; ++f) P.fs[f] = s->fs[f];
  for (int d = 0; d < 3; ++d) P.es[d] = s->es[d];
  for (int d = 0; d < 3; ++d) P.cvel[d] = s->cvel[d];
  P.Ne = s->Ne;
  P.a = a;
  P.b = b;
  P.dt = dt;
  P.mode = mode;
  return P;
}

int launch_stage(dgk_solver *s, int mode, double a, double b, double dt) {
  if (!s->structured) {
    int rc = launch_halo(s, s->q[s->cur]);
    if (rc) return rc;
  }
  StageParams P = make_params(s, mode, a, b, dt);
  s->sk.launch(s->sk, P, s->ops, s->Lift, s->stream);
  HIPOK(hipGetLastError());
  if (mode != kModeTend) {
    const int out = 3 - s->cur - s->c0;
    if (mode == kModeStage1) s->c0 = s->cur;
    s->cur = out;
  }
  return DGK_OK;
}

void drop_graphs(dgk_solver *s) {
  for (auto &g : s->graph)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
}

// Verifies that the reference tables describe a periodic structured mesh with
// axis-aligned unit normals and uniform Fscale/Escale, and if so fills the
// compressed neighbour table and constants.  Returns false otherwise.
bool compress_tables(int Nq, long long Ne, const int32_t *VMapM, const int32_t *VMapP,
                     const double *nrm, const double *Escale, const double *Fscale,
                     const int32_t *halo, int NhaloNode, std::vector<int32_t> &nbr,
                     double fs[6], double es[3]) {
  const int Np = Nq * Nq * Nq, Nfp = Nq * Nq, NfpTot = 6 * Nfp;
  const long long NfNe = (long long)NfpTot * Ne, NpNe = (long long)Np * Ne;
  const long long owned = (long long)Np * Ne;
  std::vector<int> fnode(6 * Nfp);
  for (int f = 0; f < 6; ++f)
    for (int fp = 0; fp < Nfp; ++fp) fnode[f * Nfp + fp] = face_node_host(Nq, f, fp % Nq, fp / Nq);
  nbr.assign(6 * Ne, 0);
  for (int f = 0; f < 6; ++f) fs[f] = Fscale[f * Nfp];
  for (int d = 0; d < 3; ++d) es[d] = Escale[d * NpNe];
  for (long long e = 0; e < Ne; ++e) {
    for (int d = 0; d < 3; ++d)
      for (int n = 0; n < Np; ++n)
        if (Escale[n + e * Np + d * NpNe] != es[d]) return false;
    for (int f = 0; f < 6; ++f) {
      const int opp = kFace[f].opp;
      long long eP = -1;
      for (int fp = 0; fp < Nfp; ++fp) {
        const long long t = f * Nfp + fp + e * NfpTot;
        for (int d = 0; d < 3; ++d) {
          const double want = d == kFace[f].axis ? (double)kFace[f].sign : 0.0;
          if (nrm[t + d * NfNe] != want) return false;
        }
        if (Fscale[t] != fs[f]) return false;
        if ((long long)VMapM[t] != fnode[f * Nfp + fp] + 1 + e * Np) return false;
        long long iP = (long long)VMapP[t] - 1;
        if (iP >= owned) {
          const long long ib = iP - owned;
          if (ib >= NhaloNode) return false;
          iP = (long long)halo[ib] - 1;
        }
        if (iP < 0 || iP >= owned) return false;
        const long long cand = iP / Np;
        if (iP - cand * Np != fnode[opp * Nfp + fp]) return false;
        if (eP < 0) eP = cand;
        else if (cand != eP) return false;
      }
      nbr[e * 6 + f] = (int32_t)eP;
    }
  }
  return true;
}

// When a compressed neighbour table is exactly the periodic structured mesh
// of the reference's numbering (mod_mesh.f90:486-493), records its extents
// so that the pencil kernel may walk it in tiles.  The -x, -y and -z
// neighbours of element 1 are (NeX,1,1), (1,NeY,1) and (1,1,NeZ).
void infer_structured_extents(dgk_solver *s, const std::vector<int32_t> &nbr) {
  const long long nx = (long long)nbr[3] + 1;
  if (nx < 1 || s->Ne % nx) return;
  const long long ny = nbr[0] / nx + 1;
  if (ny < 1 || (s->Ne / nx) % ny) return;
  const long long nz = s->Ne / (nx * ny);
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const long long e = i + nx * (j + ny * k);
        for (int f = 0; f < 6; ++f) {
          bool wrapped;
          if (nbr[e * 6 + f] != neighbour_elem(f, i, j, k, (int)nx, (int)ny, (int)nz, &wrapped))
            return;
        }
      }
  s->NeX = (int)nx;
  s->NeY = (int)ny;
  s->NeZ = (int)nz;
}

int finish_create(dgk_solver *s, const double *D1D, const double *Lift) {
  int rc;
  const int Nq = s->Nq, Np = s->Np;
  const size_t nq2 = (size_t)Nq * Nq, nlift = (size_t)6 * Np;
  if ((rc = dalloc(&s->D, nq2)) || (rc = dalloc(&s->Lift, nlift)) ||
      (rc = dalloc(&s->lfac, 4 * (size_t)Nq)) || (rc = dalloc(&s->Drow, nq2)) ||
      (rc = dalloc(&s->LiftP, lift_delta_doubles(Nq))) ||
      (rc = dalloc(&s->Drow32, nq2 + 3 * Nq + (Nq == 16 ? kMfmaFitDoubles : 0))))
    return rc;
  // the delta-packed Lift of the pencil kernels (without it, an exact lift
  // takes the node kernel)
  {
    std::vector<double> lp(lift_delta_doubles(Nq));
    s->lift_packed = pack_lift_deltas(Nq, Lift, lp.data());
    if (s->lift_packed)
      HIPOK(hipMemcpy(s->LiftP, lp.data(), lp.size() * sizeof(double), hipMemcpyHostToDevice));
    s->ops.LiftP = s->LiftP;
  }
  HIPOK(hipMemcpy(s->D, D1D, nq2 * sizeof(double), hipMemcpyHostToDevice));
  HIPOK(hipMemcpy(s->Lift, Lift, nlift * sizeof(double), hipMemcpyHostToDevice));
  std::vector<double> dr(nq2);
  for (int k = 0; k < Nq; ++k)
    for (int l = 0; l < Nq; ++l) dr[(size_t)k * Nq + l] = D1D[k + (size_t)l * Nq];
  HIPOK(hipMemcpy(s->Drow, dr.data(), nq2 * sizeof(double), hipMemcpyHostToDevice));
  s->ops.Drow = s->Drow;
  // [l_-, l_+], then the pairs (l_-(n), l_+(n)) that the one-wavefront stage
  // reads with one 16-byte load per index
  std::vector<double> lf(4 * (size_t)Nq);
  for (int i = 0; i < Nq; ++i) {
    lf[i] = Lift[i + (size_t)Np * 3];       // face 4 (-x): l_-(i) = Lift(i,1,1,4)
    lf[Nq + i] = Lift[i + (size_t)Np * 1];  // face 2 (+x): l_+(i) = Lift(i,1,1,2)
  }
  for (int i = 0; i < Nq; ++i) {
    lf[2 * Nq + 2 * i] = lf[i];
    lf[2 * Nq + 2 * i + 1] = lf[Nq + i];
  }
  HIPOK(hipMemcpy(s->lfac, lf.data(), lf.size() * sizeof(double), hipMemcpyHostToDevice));
  s->ops.lfac = s->lfac;
  // FP32 split copies (x = hi + lo, hi = fl32(x), lo = fl32(x - hi)) and the
  // node tables of the line fit; at Nq = 16 the matrix-core form's fit
  // tables after them
  {
    std::vector<double> d32(nq2 + 3 * Nq + (Nq == 16 ? kMfmaFitDoubles : 0));
    split_tables(Nq, dr.data(), d32.data());
    if (Nq == 16) mfma_fit_tables(dr.data(), d32.data() + nq2 + 3 * Nq);
    HIPOK(hipMemcpy(s->Drow32, d32.data(), d32.size() * sizeof(double), hipMemcpyHostToDevice));
    s->ops.Drow32 = s->Drow32;
    s->ops.mfit = Nq == 16 ? s->Drow32 + nq2 + 3 * Nq : nullptr;
  }
  if ((rc = alloc_fields(s))) return rc;
  if ((rc = configure_launch(s))) return rc;
  HIPOK(hipStreamSynchronize(s->stream));
  return DGK_OK;
}

template <class T>
int upload(T **dst, const T *src, size_t n, hipStream_t st) {
  int rc = dalloc(dst, n);
  if (rc) return rc;
  if (n) HIPOK(hipMemcpyAsync(*dst, src, n * sizeof(T), hipMemcpyHostToDevice, st));
  HIPOK(hipStreamSynchronize(st));
  return DGK_OK;
}

}  // namespace

extern "C" {

int dgk_destroy(dgk_solver *s) {
  if (!s) return DGK_OK;
  drop_graphs(s);
  void *bufs[] = {s->q[0], s->q[1], s->q[2], s->u, s->v, s->w, s->dqdt, s->D, s->Lift, s->lfac, s->Drow, s->LiftP,
                  s->nbr, s->vmapM, s->vmapP, s->halo, s->nrm, s->fscale, s->escale,
                  s->part, s->mm, s->diag, s->Drow32, s->tile_order};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
  return DGK_OK;
}

int dgk_create(dgk_solver **out, int Nq, int Ne, int NeA, int NfpTot, const double *D1D,
               const double *Lift_mat, const int32_t *VMapM, const int32_t *VMapP,
               const double *normal_fn, const double *Escale, const double *Fscale,
               const int32_t *halo_src_map, int NhaloNode, int typeid_, int flags) {
  clear_error();
  if (!out) return fail(DGK_ERR_ARG, "null solver pointer");
  *out = nullptr;
  if (Nq < 2 || Nq > 16) return fail(DGK_ERR_ORDER, "PolyOrder %d outside 1..15", Nq - 1);
  if (NfpTot != 6 * Nq * Nq || Ne < 1 || NeA < Ne || NhaloNode < 0)
    return fail(DGK_ERR_ARG, "inconsistent sizes Nq=%d NfpTot=%d Ne=%d NeA=%d", Nq, NfpTot, Ne,
                NeA);
  const long long Np = (long long)Nq * Nq * Nq;
  if (NhaloNode > Np * (NeA - Ne)) return fail(DGK_ERR_ARG, "NhaloNode exceeds the halo slack");
  if (!D1D || !Lift_mat || !VMapM || !VMapP || !normal_fn || !Escale || !Fscale ||
      (NhaloNode && !halo_src_map))
    return fail(DGK_ERR_ARG, "null table");
  const int order = order_of(typeid_);
  if (order < 0) return fail(DGK_ERR_OPTTYPE, "typeid %d is not GENERAL, OPT1 or HIP", typeid_);
  // Validate index ranges before anything reaches the device.
  const long long NfNe = (long long)NfpTot * Ne, total = Np * NeA;
  for (long long t = 0; t < NfNe; ++t)
    if (VMapM[t] < 1 || VMapM[t] > total || VMapP[t] < 1 || VMapP[t] > total)
      return fail(DGK_ERR_ARG, "VMapM/VMapP entry %lld out of 1..Np*NeA", t + 1);
  for (int ib = 0; ib < NhaloNode; ++ib)
    if (halo_src_map[ib] < 1 || halo_src_map[ib] > Np * Ne)
      return fail(DGK_ERR_ARG, "halo_src_map entry %d out of the owned range", ib + 1);

  dgk_solver *s = new dgk_solver;
  s->Nq = Nq;
  s->Np = (int)Np;
  s->Nfp = Nq * Nq;
  s->NfpTot = NfpTot;
  s->Ne = Ne;
  s->NeA = NeA;
  s->NhaloNode = NhaloNode;
  s->order = (flags & DGK_FLAG_MIXED) ? kMixed : order;
  s->fma = (flags & DGK_FLAG_FMA) != 0;
  // the mixed form lifts with the 1D vectors
  s->flags = (flags & (DGK_FLAG_MIXED | DGK_FLAG_MFMA)) ? (flags | DGK_FLAG_FACTORED_LIFT) : flags;
  int rc = DGK_OK;
  if (hipGetDevice(&s->device) != hipSuccess ||
      hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
    delete s;
    return fail(DGK_ERR_HIP, "hipStreamCreate failed");
  }
  std::vector<int32_t> nbr;
  s->structured = !(flags & DGK_FLAG_FORCE_TABLES) &&
                  compress_tables(Nq, Ne, VMapM, VMapP, normal_fn, Escale, Fscale, halo_src_map,
                                  NhaloNode, nbr, s->fs, s->es);
  if (s->structured) {
    infer_structured_extents(s, nbr);
    rc = upload(&s->nbr, nbr.data(), nbr.size(), s->stream);
  } else {
    if (!rc) rc = upload(&s->vmapM, VMapM, (size_t)NfNe, s->stream);
    if (!rc) rc = upload(&s->vmapP, VMapP, (size_t)NfNe, s->stream);
    if (!rc) rc = upload(&s->nrm, normal_fn, (size_t)(3 * NfNe), s->stream);
    if (!rc) rc = upload(&s->fscale, Fscale, (size_t)NfNe, s->stream);
    if (!rc) rc = upload(&s->escale, Escale, (size_t)(3 * Np * Ne), s->stream);
  }
  if (!rc) rc = upload(&s->halo, halo_src_map, (size_t)NhaloNode, s->stream);
  if (!rc) rc = finish_create(s, D1D, Lift_mat);
  if (rc) {
    dgk_destroy(s);
    return rc;
  }
  *out = s;
  return DGK_OK;
}

int dgk_create_structured(dgk_solver **out, int NeX, int NeY, int NeZ, int p, double domx,
                          double domy, double domz, const double *D1D, const double *Lift_mat,
                          int typeid_, int flags) {
  clear_error();
  if (!out) return fail(DGK_ERR_ARG, "null solver pointer");
  *out = nullptr;
  int32_t sz[7];
  int rc = dgk_mesh_sizes(NeX, NeY, NeZ, p, sz);
  if (rc) return rc;
  if (flags & DGK_FLAG_FORCE_TABLES)
    return fail(DGK_ERR_ARG, "DGK_FLAG_FORCE_TABLES needs dgk_create with tables");
  if (!D1D || !Lift_mat) return fail(DGK_ERR_ARG, "null operator");
  const int order = order_of(typeid_);
  if (order < 0) return fail(DGK_ERR_OPTTYPE, "typeid %d is not GENERAL, OPT1 or HIP", typeid_);
  dgk_solver *s = new dgk_solver;
  s->Nq = sz[0];
  s->Np = sz[1];
  s->Nfp = sz[2];
  s->NfpTot = sz[3];
  s->Ne = sz[4];
  s->NeA = sz[5];
  s->NhaloNode = sz[6];
  s->order = (flags & DGK_FLAG_MIXED) ? kMixed : order;
  s->fma = (flags & DGK_FLAG_FMA) != 0;
  // the mixed form lifts with the 1D vectors
  s->flags = (flags & (DGK_FLAG_MIXED | DGK_FLAG_MFMA)) ? (flags | DGK_FLAG_FACTORED_LIFT) : flags;
  s->structured = true;
  s->NeX = NeX;
  s->NeY = NeY;
  s->NeZ = NeZ;
  if (hipGetDevice(&s->device) != hipSuccess ||
      hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
    delete s;
    return fail(DGK_ERR_HIP, "hipStreamCreate failed");
  }
  const double dl[3] = {domx / (double)NeX, domy / (double)NeY, domz / (double)NeZ};
  for (int f = 0; f < 6; ++f) s->fs[f] = 2.0 / dl[kFace[f].axis];  // mod_mesh.f90:366-394
  for (int d = 0; d < 3; ++d) s->es[d] = 2.0 / dl[d];               // mod_mesh.f90:147-149
  std::vector<int32_t> nbr(6 * (size_t)s->Ne);
  for (int k = 0; k < NeZ; ++k)
    for (int j = 0; j < NeY; ++j)
      for (int i = 0; i < NeX; ++i) {
        const long long e = i + (long long)j * NeX + (long long)k * NeX * NeY;
        for (int f = 0; f < 6; ++f) {
          bool wrapped;
          nbr[e * 6 + f] = (int32_t)neighbour_elem(f, i, j, k, NeX, NeY, NeZ, &wrapped);
        }
      }
  std::vector<int32_t> halo(s->NhaloNode);
  rc = dgk_mesh_setup(NeX, NeY, NeZ, p, domx, domy, domz, nullptr, nullptr, nullptr, nullptr,
                      nullptr, nullptr, nullptr, halo.data());
  if (!rc) rc = upload(&s->nbr, nbr.data(), nbr.size(), s->stream);
  if (!rc) rc = upload(&s->halo, halo.data(), halo.size(), s->stream);
  if (!rc) rc = finish_create(s, D1D, Lift_mat);
  if (rc) {
    dgk_destroy(s);
    return rc;
  }
  *out = s;
  return DGK_OK;
}

int dgk_set_fields(dgk_solver *s, const double *q, const double *u, const double *v,
                   const double *w) {
  if (!s || !q || !u || !v || !w) return fail(DGK_ERR_ARG, "null argument");
  const size_t nb = (size_t)s->Np * s->NeA * sizeof(double);
  if (s->flags & DGK_FLAG_CONST_VELOCITY) {
    // the compressed form streams no velocity: every owned node must carry
    // the same (u,v,w)
    const double *vel[3] = {u, v, w};
    const size_t owned = (size_t)s->Np * s->Ne;
    for (int d = 0; d < 3; ++d) {
      size_t n = 0;
      if (owned > 1 && find_not_equal(vel[d] + 1, owned - 1, vel[d][0], &n))
        return fail(DGK_ERR_ARG, "DGK_FLAG_CONST_VELOCITY: velocity component %d is not "
                                 "uniform (node %zu)", d + 1, n + 2);
      s->cvel[d] = vel[d][0];
    }
    drop_graphs(s);  // the velocity is a launch argument
  }
  s->cur = 0;
  s->c0 = 2;
  int rc;
  if ((rc = copy_h2d(s->q[0], q, nb, s->stream))) return rc;
  if ((rc = launch_halo(s, s->q[0]))) return rc;
  double *vel_dev[3] = {s->u, s->v, s->w};
  const double *vel_host[3] = {u, v, w};
  for (int d = 0; d < 3; ++d) {
    if (s->flags & DGK_FLAG_CONST_VELOCITY) {
      // verified uniform above: the device copy (halo included) is the
      // constant, written on the device instead of uploaded
      const long long n = (long long)s->Np * s->NeA;
      const unsigned grid = (unsigned)std::min<long long>((n + 255) / 256, 4096);
      hipLaunchKernelGGL(fill_kernel, dim3(grid), dim3(256), 0, s->stream, vel_dev[d], n,
                         s->cvel[d]);
      HIPOK(hipGetLastError());
    } else {
      if ((rc = copy_h2d(vel_dev[d], vel_host[d], nb, s->stream))) return rc;
      if ((rc = launch_halo(s, vel_dev[d]))) return rc;
    }
  }
  HIPOK(hipStreamSynchronize(s->stream));
  s->fields_set = true;
  return DGK_OK;
}

int dgk_field_ptr(dgk_solver *s, int which, double **dptr) {
  if (!s || !dptr) return fail(DGK_ERR_ARG, "null argument");
  switch (which) {
  case DGK_FIELD_Q: *dptr = s->q[s->cur]; break;
  case DGK_FIELD_Q0: *dptr = s->q[s->c0]; break;
  case DGK_FIELD_DQDT: *dptr = s->dqdt; break;
  case DGK_FIELD_U: *dptr = s->u; break;
  case DGK_FIELD_V: *dptr = s->v; break;
  case DGK_FIELD_W: *dptr = s->w; break;
  default: return fail(DGK_ERR_ARG, "unknown field %d", which);
  }
  return DGK_OK;
}

int dgk_get_field(dgk_solver *s, int which, double *out) {
  double *src = nullptr;
  int rc = dgk_field_ptr(s, which, &src);
  if (rc) return rc;
  if (!out) return fail(DGK_ERR_ARG, "null output");
  if (which == DGK_FIELD_Q && (rc = launch_halo(s, src))) return rc;
  return copy_d2h(out, src, (size_t)s->Np * s->NeA * sizeof(double), s->stream);
}

int dgk_cal_tend(dgk_solver *s) {
  if (!s) return fail(DGK_ERR_ARG, "null solver");
  if (!s->fields_set) return fail(DGK_ERR_STATE, "dgk_set_fields has not been called");
  return launch_stage(s, kModeTend, 0.0, 0.0, 0.0);
}

int dgk_rk_stage(dgk_solver *s, int stage, double dt) {
  if (!s) return fail(DGK_ERR_ARG, "null solver");
  if (!s->fields_set) return fail(DGK_ERR_STATE, "dgk_set_fields has not been called");
  if (stage < 1 || stage > 3) return fail(DGK_ERR_ARG, "stage %d outside 1..3", stage);
  return launch_stage(s, stage == 1 ? kModeStage1 : kModeStageN, kRkA[stage - 1],
                      kRkB[stage - 1], dt);
}

int dgk_step(dgk_solver *s, double dt) {
  if (!s) return fail(DGK_ERR_ARG, "null solver");
  if (!s->fields_set) return fail(DGK_ERR_STATE, "dgk_set_fields has not been called");
  if (s->graph_dt != dt) drop_graphs(s);
  // The buffers a step reads and writes depend on where q and q0 sit when it
  // starts: one graph per starting (cur, c0).
  const int key = 3 * s->cur + s->c0, cur0 = s->cur, c00 = s->c0;
  if (!s->graph[key]) {
    hipGraph_t g;
    HIPOK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    int rc = DGK_OK;
    for (int st = 1; st <= 3 && !rc; ++st)
      rc = launch_stage(s, st == 1 ? kModeStage1 : kModeStageN, kRkA[st - 1], kRkB[st - 1], dt);
    hipError_t ce = hipStreamEndCapture(s->stream, &g);
    s->graph_next[key][0] = s->cur;
    s->graph_next[key][1] = s->c0;
    s->cur = cur0;
    s->c0 = c00;
    if (rc) return rc;
    if (ce != hipSuccess) return fail(DGK_ERR_HIP, "graph capture: %s", hipGetErrorString(ce));
    hipError_t ie = hipGraphInstantiate(&s->graph[key], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ie != hipSuccess) return fail(DGK_ERR_HIP, "graph instantiate: %s", hipGetErrorString(ie));
    s->graph_dt = dt;
  }
  HIPOK(hipGraphLaunch(s->graph[key], s->stream));
  s->cur = s->graph_next[key][0];
  s->c0 = s->graph_next[key][1];
  return DGK_OK;
}

int dgk_run(dgk_solver *s, double dt, int nstep, double *minmax) {
  if (!s) return fail(DGK_ERR_ARG, "null solver");
  if (nstep < 0) return fail(DGK_ERR_ARG, "negative nstep");
  if (minmax && nstep > s->mm_cap) {
    if (s->mm) (void)hipFree(s->mm);
    s->mm = nullptr;
    int rc = dalloc(&s->mm, 2 * (size_t)nstep);
    if (rc) return rc;
    s->mm_cap = nstep;
  }
  if (!s->fields_set) return fail(DGK_ERR_STATE, "dgk_set_fields has not been called");
  const long long owned = (long long)s->Np * s->Ne;
  const int nb = (int)std::min<long long>(1024, (owned + 255) / 256);
  // The stages are launched on the stream one after the other, not replayed
  // from the captured step (dgk_step): a replayed graph leaves about 4 us
  // between its kernel nodes, stream launches of these kernels run back to
  // back (8^3 p=3: 3.8 instead of 5.4 us per stage; the same at 64^3 p=7).
  for (int it = 0; it < nstep; ++it) {
    int rc = DGK_OK;
    for (int st = 1; st <= 3 && !rc; ++st)
      rc = launch_stage(s, st == 1 ? kModeStage1 : kModeStageN, kRkA[st - 1], kRkB[st - 1], dt);
    if (rc) return rc;
    if (minmax) {
      hipLaunchKernelGGL(minmax_partial_kernel, dim3(nb), dim3(256), 0, s->stream,
                         s->q[s->cur], owned, s->part);
      hipLaunchKernelGGL(minmax_final_kernel, dim3(1), dim3(256), 0, s->stream, s->part, nb,
                         s->mm + 2 * it);
      HIPOK(hipGetLastError());
    }
  }
  if (minmax && nstep)
    HIPOK(hipMemcpyAsync(minmax, s->mm, 2 * (size_t)nstep * sizeof(double),
                         hipMemcpyDeviceToHost, s->stream));
  HIPOK(hipStreamSynchronize(s->stream));
  if (minmax)
    for (int it = 0; it < nstep; ++it)
      if (!std::isfinite(minmax[2 * it]) || !std::isfinite(minmax[2 * it + 1]))
        return fail(DGK_ERR_NONFINITE, "the state is not finite after step %d (min %g, max %g)",
                    it + 1, minmax[2 * it], minmax[2 * it + 1]);
  return DGK_OK;
}

int dgk_synchronize(dgk_solver *s) {
  if (!s) return fail(DGK_ERR_ARG, "null solver");
  HIPOK(hipStreamSynchronize(s->stream));
  return DGK_OK;
}

int dgk_query(dgk_solver *s, int what, int64_t *value) {
  if (!s || !value) return fail(DGK_ERR_ARG, "null argument");
  switch (what) {
  case 0: *value = s->structured ? 1 : 0; break;
  case 1: *value = s->Ne; break;
  case 2: *value = s->Np; break;
  case 3: *value = s->NeA; break;
  case 4: *value = s->sk.grid; break;
  case 5: *value = s->sk.nt; break;
  case 6: *value = s->sk.pencil ? 1 : 0; break;
  case 7: *value = (int64_t)s->sk.lds; break;
  case 8: *value = current_variant(s); break;
  case 9: *value = tiled_order(s) ? 1 : 0; break;
  default: return fail(DGK_ERR_ARG, "unknown query %d", what);
  }
  return DGK_OK;
}

int dgk_num_variants(dgk_solver *s, int *n) {
  if (!s || !n) return fail(DGK_ERR_ARG, "null argument");
  *n = pencil_variants(s);
  return DGK_OK;
}

int dgk_set_variant(dgk_solver *s, int variant) {
  if (!s) return fail(DGK_ERR_ARG, "null solver");
  if (!s->sk.pencil) {
    if (variant != 0) return fail(DGK_ERR_ARG, "this stage kernel has one variant");
    return DGK_OK;
  }
  if (variant < -1 || variant >= pencil_variants(s))
    return fail(DGK_ERR_ARG, "variant %d outside -1..%d", variant, pencil_variants(s) - 1);
  // configure_launch fills s->sk (and the tile table) for s->variant: on
  // failure the previous variant and its launch descriptor are put back
  const int old_variant = s->variant;
  const StageKernel old_sk = s->sk;
  s->variant = variant;
  const int rc = configure_launch(s);
  if (rc || !s->sk.launch || (tiled_order(s) && !s->tile_order)) {
    s->variant = old_variant;
    s->sk = old_sk;
    return rc ? rc : fail(DGK_ERR_STATE, "variant %d has no stage kernel", variant);
  }
  drop_graphs(s);
  return DGK_OK;
}

int dgk_diagnostics(dgk_solver *s, const double *w1D, const double *ref, double *out) {
  if (!s || !w1D || !out) return fail(DGK_ERR_ARG, "bad argument");
  if (!s->fields_set) return fail(DGK_ERR_STATE, "dgk_set_fields has not been called");
  if (!s->diag) {
    int rc = dalloc(&s->diag, (size_t)kDiagScratch);
    if (rc) return rc;
  }
  const double ones[3] = {1.0, 1.0, 1.0};
  return device_diagnostics(s->q[s->cur], ref, (long long)s->Np * s->Ne, s->Nq,
                            s->nbr ? s->es : ones, s->nbr ? nullptr : s->escale, w1D, s->diag,
                            s->stream, out);
}

int dgk_diagnostics_noalloc(dgk_solver *s, const double *w1D, const double *ref,
                            double *out_dev) {
  if (!s || !w1D || !out_dev) return fail(DGK_ERR_ARG, "bad argument");
  if (!s->fields_set) return fail(DGK_ERR_STATE, "dgk_set_fields has not been called");
  double *part = nullptr;
  int rc = noalloc_scratch(s->device, &part);
  if (rc) return rc;
  const double ones[3] = {1.0, 1.0, 1.0};
  return diagnostics_launch(s->q[s->cur], ref, (long long)s->Np * s->Ne, s->Nq,
                            s->nbr ? s->es : ones, s->nbr ? nullptr : s->escale, w1D, part,
                            kNoAllocBlocks, s->stream, out_dev);
}

int dgk_autotune(dgk_solver *s, double dt, int nrep, int *variant_out, double *ms_out) {
  if (!s || nrep < 1) return fail(DGK_ERR_ARG, "bad argument");
  if (!s->fields_set) return fail(DGK_ERR_STATE, "dgk_set_fields has not been called");
  const int nvar = pencil_variants(s);
  const int nstage = (nrep + 2) / 3 * 3;  // whole steps
  // the trial stages advance the state: keep q and q0 aside and restore them
  HIPOK(hipStreamSynchronize(s->stream));
  const size_t bytes = (size_t)s->Np * (size_t)s->NeA * sizeof(double);
  double *save = nullptr;
  int rc = dalloc(&save, 2 * (size_t)s->Np * (size_t)s->NeA);
  if (rc) return rc;
  const int cur0 = s->cur, c00 = s->c0, var0 = s->variant;
  int best = -1;
  double best_ms = 0.0;
  if (hipMemcpy(save, s->q[cur0], bytes, hipMemcpyDeviceToDevice) != hipSuccess ||
      hipMemcpy(save + bytes / sizeof(double), s->q[c00], bytes, hipMemcpyDeviceToDevice) !=
          hipSuccess)
    rc = fail(DGK_ERR_HIP, "autotune: state copy failed");
  for (int v = 0; v < nvar && !rc; ++v) {
    double ms = 0.0;
    if (nvar > 1 && (rc = dgk_set_variant(s, v))) break;
    if ((rc = dgk_time_stages(s, dt, 3, &ms))) break;  // warm-up
    if ((rc = dgk_time_stages(s, dt, nstage, &ms))) break;
    if (best < 0 || ms < best_ms) {
      best = v;
      best_ms = ms;
    }
  }
  s->cur = cur0;
  s->c0 = c00;
  if (hipMemcpy(s->q[cur0], save, bytes, hipMemcpyDeviceToDevice) != hipSuccess ||
      hipMemcpy(s->q[c00], save + bytes / sizeof(double), bytes, hipMemcpyDeviceToDevice) !=
          hipSuccess) {
    if (!rc) rc = fail(DGK_ERR_HIP, "autotune: state restore failed");
  }
  (void)hipFree(save);
  if (rc) {
    if (nvar > 1) (void)dgk_set_variant(s, var0);
    return rc;
  }
  if (nvar > 1 && (rc = dgk_set_variant(s, best))) return rc;
  if (variant_out) *variant_out = nvar > 1 ? best : 0;
  if (ms_out) *ms_out = best_ms;
  return DGK_OK;
}

int dgk_time_stages(dgk_solver *s, double dt, int nstage, double *ms_per_stage) {
  if (!s || !ms_per_stage || nstage < 1) return fail(DGK_ERR_ARG, "bad argument");
  if (nstage % 3)
    return fail(DGK_ERR_ARG, "nstage %d is not a whole number of SSP-RK3 steps (3 stages each)",
                nstage);
  if (!s->fields_set) return fail(DGK_ERR_STATE, "dgk_set_fields has not been called");
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = DGK_OK;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess ||
      hipEventRecord(e0, s->stream) != hipSuccess)
    rc = fail(DGK_ERR_HIP, "time_stages: event setup failed");
  for (int it = 0; it < nstage && !rc; ++it) {
    const int st = it % 3;
    rc = launch_stage(s, st == 0 ? kModeStage1 : kModeStageN, kRkA[st], kRkB[st], dt);
  }
  float ms = 0.f;
  if (!rc && (hipEventRecord(e1, s->stream) != hipSuccess ||
              hipEventSynchronize(e1) != hipSuccess ||
              hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
    rc = fail(DGK_ERR_HIP, "time_stages: event timing failed");
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (rc) return rc;
  *ms_per_stage = ms / nstage;
  return DGK_OK;
}

int dgk_advect3d_eq_cal_tend(double *dqdt, const double *q, const double *u, const double *v,
                             const double *w, const double *D1D, const double *D1D_tr,
                             const double *Lift_mat, const int32_t *VMapM,
                             const int32_t *VMapP, const double *normal_fn,
                             const double *Escale, const double *Fscale, int Nq, int Np,
                             int NfpTot, int Ne, int NeA, int typeid_) {
  clear_error();
  if (!dqdt || !q || !u || !v || !w) return fail(DGK_ERR_ARG, "null field");
  if (Np != Nq * Nq * Nq) return fail(DGK_ERR_ARG, "Np != Nq^3");
  (void)D1D_tr;  // the transpose is implied by D1D (mod_mesh.f90:263)
  dgk_solver *s = nullptr;
  // The caller's q,u,v,w carry their halos already, so no halo map is needed:
  // the table path gathers through VMapP directly.
  int rc = dgk_create(&s, Nq, Ne, NeA, NfpTot, D1D, Lift_mat, VMapM, VMapP, normal_fn, Escale,
                      Fscale, nullptr, 0, typeid_ == DGK_TYPE_GENERAL ? DGK_TYPE_GENERAL
                                                                       : DGK_TYPE_OPT1,
                      DGK_FLAG_FORCE_TABLES);
  if (rc) return rc;
  const size_t nb = (size_t)Np * NeA * sizeof(double);
  s->cur = 0;
  rc = DGK_OK;
  double *dev[4] = {s->q[0], s->u, s->v, s->w};
  const double *host[4] = {q, u, v, w};
  for (int b = 0; b < 4 && !rc; ++b) rc = copy_h2d(dev[b], host[b], nb, s->stream);
  s->fields_set = true;
  if (!rc) rc = launch_stage(s, kModeTend, 0.0, 0.0, 0.0);
  if (!rc && hipStreamSynchronize(s->stream) != hipSuccess)
    rc = fail(DGK_ERR_HIP, "tendency kernel failed");
  // dqdt(:,1:Ne) only, as cal_dqdt writes (mod_advect3d_eq.f90:194)
  if (!rc) rc = copy_d2h(dqdt, s->dqdt, (size_t)Np * Ne * sizeof(double), s->stream);
  dgk_destroy(s);
  return rc;
}

int dgk_divlike_dirXYZ_batched(int Nq, int nelem, int typeid_, const double *Mat,
                               const double *Mat_tr, const double *in_x, const double *in_y,
                               const double *in_z, double *out_x, double *out_y, double *out_z,
                               void *stream) {
  clear_error();
  if (Nq < 2 || Nq > 16) return fail(DGK_ERR_ORDER, "PolyOrder %d outside 1..15", Nq - 1);
  if (nelem < 0) return fail(DGK_ERR_ARG, "negative nelem");
  if (nelem == 0) return DGK_OK;
  const int order = order_of(typeid_);
  if (order < 0) return fail(DGK_ERR_OPTTYPE, "typeid %d", typeid_);
  if (!Mat || !Mat_tr || !in_x || !in_y || !in_z || !out_x || !out_y || !out_z)
    return fail(DGK_ERR_ARG, "null pointer");
  const size_t lds = sizeof(double) * (3 * (size_t)Nq * Nq * Nq + 2 * (size_t)Nq * Nq);
  hipStream_t st = (hipStream_t)stream;
  using Fn = void (*)(const double *, const double *, const double *, const double *,
                      const double *, double *, double *, double *);
  Fn fn = nullptr;
#define DGK_DIV_CASE(NQ)                                                                   \
  case NQ:                                                                                 \
    fn = order == kGeneral ? divlike_batched_kernel<NQ, kGeneral> : divlike_batched_kernel<NQ, kOpt1>; \
    break;
  switch (Nq) {
    DGK_DIV_CASE(2) DGK_DIV_CASE(3) DGK_DIV_CASE(4) DGK_DIV_CASE(5) DGK_DIV_CASE(6)
    DGK_DIV_CASE(7) DGK_DIV_CASE(8) DGK_DIV_CASE(9) DGK_DIV_CASE(10) DGK_DIV_CASE(11)
    DGK_DIV_CASE(12) DGK_DIV_CASE(13) DGK_DIV_CASE(14) DGK_DIV_CASE(15) DGK_DIV_CASE(16)
  }
#undef DGK_DIV_CASE
  HIPOK(hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lds));
  hipLaunchKernelGGL(fn, dim3(nelem), dim3(256), lds, st, Mat, Mat_tr, in_x, in_y, in_z, out_x,
                     out_y, out_z);
  HIPOK(hipGetLastError());
  return DGK_OK;
}

int dgk_lift_hexahedral_batched(int Nq, int nelem, const double *Lift, const double *in,
                                double *out, void *stream) {
  clear_error();
  if (Nq < 2 || Nq > 16) return fail(DGK_ERR_ORDER, "PolyOrder %d outside 1..15", Nq - 1);
  if (nelem < 0) return fail(DGK_ERR_ARG, "negative nelem");
  if (nelem == 0) return DGK_OK;
  if (!Lift || !in || !out) return fail(DGK_ERR_ARG, "null pointer");
  hipStream_t st = (hipStream_t)stream;
  using Fn = void (*)(const double *, const double *, double *, long long);
  Fn fn = nullptr;
  int nt = 0, epb = 1;
  size_t lds = 0;
#define DGK_LIFT_CASE(NQ)                                                      \
  case NQ:                                                                     \
    fn = lift_batched_kernel<NQ>;                                              \
    nt = LiftShape<NQ>::NT;                                                    \
    epb = LiftShape<NQ>::EPB;                                                  \
    lds = LiftShape<NQ>::LDS;                                                  \
    break;
  switch (Nq) {
    DGK_LIFT_CASE(2) DGK_LIFT_CASE(3) DGK_LIFT_CASE(4) DGK_LIFT_CASE(5) DGK_LIFT_CASE(6)
    DGK_LIFT_CASE(7) DGK_LIFT_CASE(8) DGK_LIFT_CASE(9) DGK_LIFT_CASE(10) DGK_LIFT_CASE(11)
    DGK_LIFT_CASE(12) DGK_LIFT_CASE(13) DGK_LIFT_CASE(14) DGK_LIFT_CASE(15) DGK_LIFT_CASE(16)
  }
#undef DGK_LIFT_CASE
  // enough workgroups to fill the chip several times over; each then walks
  // the elements in a grid stride
  const long long groups = (nelem + epb - 1) / epb;
  const unsigned grid = (unsigned)std::min<long long>(groups, 4096);
  hipLaunchKernelGGL(fn, dim3(grid), dim3(nt), lds, st, Lift, in, out, (long long)nelem);
  HIPOK(hipGetLastError());
  return DGK_OK;
}

}  // extern "C"
