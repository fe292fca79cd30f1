// Batched ensemble: many independent periodic meshes of uneven polynomial
// order, size and velocity advanced together, one persistent launch per RK
// stage and the three stages of a time step captured in one hipGraph.
//
// Each instance keeps the reference layout q(Np,Ne) inside pooled device
// arrays (int64 offsets).  Work items (instance, first element) are grouped
// into classes of equal order; the persistent grid is split between the
// classes (in units of eight workgroups, one per XCD) and every workgroup
// runs items of one class only, so that it is compiled and scheduled as one
// straight code path.  A workgroup of 192 threads (three wavefronts, each on
// its own SIMD) runs an item as the pencil stage (one lane per node line) of
// up to as many elements as fill it: 48, 12, 5, 3 or 1 for Nq = 2, 4, 6, 8,
// 12.  An ensemble with an instance of Nq 14..16 (196..256 lines per
// element) uses workgroups of 256 threads instead; any ensemble may.
// Within a class the items are split into eight contiguous ranges, one per
// XCD, so that an instance's neighbouring elements stay in one L2.  The
// split of the grid between the classes follows the DOF count of each
// class, or, once dgk_ensemble_tune has run, the tile shape (workgroup width,
// elements per item of each class) and split that measured fastest.
#include "dgk_internal.hpp"
#include "dgk_kernels.hpp"
#include "dgk_pencil.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <vector>

using namespace dgk;

#define HIPOK(expr)                                                            \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess)                                                      \
      return fail(DGK_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

namespace {

constexpr int kEnsNarrow = 192;  // Nq 2..13
constexpr int kEnsWide = 256;    // Nq 2..16

struct EnsInst {
  int nq, op;            // Nq and operator slot
  long long ne;          // elements
  long long off;         // field offset (doubles) in the pools
  long long nbr_off;     // neighbour-table offset (int32)
  double es[3], fs[6], cvel[3];
  double dt;
};

struct EnsOps {          // per distinct Nq
  const double *D, *Lift, *Drow, *LiftP, *lfac;
  const double *Drow32;  // DGK_FLAG_MIXED: FP32 split rows and line-fit tables (split_tables)
};

constexpr int kEnsMaxClass = 15;  // Nq 2..16

struct EnsArgs {
  const EnsInst *inst;
  const EnsOps *ops;
  const int2 *items;     // (instance, first element), grouped by class
  int ncls;
  int cls_nq[kEnsMaxClass];
  int cls_epb[kEnsMaxClass];        // elements per item of each class
  int cls_block[kEnsMaxClass + 1];  // first workgroup of each class (multiples of 8)
  int cls_item[kEnsMaxClass + 1];   // first item of each class
  const StageParams *params;        // per instance, for this launch's stage and placement
};

bool ens_supported(int nq) { return nq >= 2 && nq <= 16; }
// workgroup size for an ensemble whose largest order has max_nq nodes a line
int ens_threads(int max_nq) { return max_nq <= 13 ? kEnsNarrow : kEnsWide; }

// Most elements per item (one workgroup): Nq^2 lanes per element, 24, 12, 5, 3
// and 1 elements for Nq 2, 4, 6, 8, 12.
int ens_item_elems(int nq, int nt) { return nt / (nq * nq); }

// LDS of an item (the ensemble stages q in LDS whatever the velocity)
template <int NQ>
size_t ens_lds_nq() {
  return sizeof(double) * PencilShape<NQ>::lean_elem(false);
}

size_t ens_elem_lds(int nq) {
  switch (nq) {
  case 2: return ens_lds_nq<2>();
  case 3: return ens_lds_nq<3>();
  case 4: return ens_lds_nq<4>();
  case 5: return ens_lds_nq<5>();
  case 6: return ens_lds_nq<6>();
  case 7: return ens_lds_nq<7>();
  case 8: return ens_lds_nq<8>();
  case 9: return ens_lds_nq<9>();
  case 10: return ens_lds_nq<10>();
  case 11: return ens_lds_nq<11>();
  case 12: return ens_lds_nq<12>();
  case 13: return ens_lds_nq<13>();
  case 14: return ens_lds_nq<14>();
  case 15: return ens_lds_nq<15>();
  default: return ens_lds_nq<16>();
  }
}

// The per-instance parameter blocks, operator slots and work items are read
// through the constant address space: scalar loads at each use, which the
// compiler may repeat instead of holding the values in registers across the
// whole stage.
typedef __attribute__((address_space(4))) const StageParams CParams;
typedef __attribute__((address_space(4))) const EnsInst CInst;
typedef __attribute__((address_space(4))) const EnsOps COps;
typedef __attribute__((address_space(4))) const int CInt;

// The items [first, first+count) of one class on nb workgroups (lb = this
// workgroup's index among them; lb % 8 is its XCD).  An item is the lean
// pencil stage of the elements first .. first+epb-1 of an instance (epb at
// most EPB, the elements whose lines fill the workgroup); lanes beyond
// epb*Nq^2 only join the barriers.
template <int NT, int NQ, int ORDER, bool FMA, bool FACT, bool CV>
__device__ __forceinline__ void ens_items(const EnsArgs &A, int first, int count, int epb, int lb,
                                          int nb, double *smem) {
  constexpr int NFP = NQ * NQ, EPB = NT / NFP;
  static_assert(EPB >= 1, "an element's lines must fit the workgroup");
  constexpr int PER = PencilShape<NQ>::lean_elem(false);  // q staged in LDS
  const int xcd = lb % kNumXcd, slot = lb / kNumXcd, nslot = nb / kNumXcd;
  const int per = (count + kNumXcd - 1) / kNumXcd;
  const int end = first + min(count, (xcd + 1) * per);
  for (int it = first + xcd * per + slot; it < end; it += nslot) {
    const int inst = ((CInt *)A.items)[2 * it], first_el = ((CInt *)A.items)[2 * it + 1];
    CParams &P = ((CParams *)A.params)[inst];
    COps &O = ((COps *)A.ops)[((CInst *)A.inst)[inst].op];
    // the lane's line and element, made opaque per item so that the address
    // arithmetic derived from them is not hoisted out of the loop and held
    // in registers through every item
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    const bool act = tid < epb * NFP;
    const int el = act ? tid / NFP : 0;
    const int pl = act ? tid - el * NFP : 0;
    long long e = first_el + el;
    if (!act || e >= P.Ne) e = -1;
    pencil_element<NQ, ORDER, FMA, FACT, true, CV, true, CParams, false, true>(
        P, ORDER == kMixed ? O.Drow32 : O.Drow, O.LiftP, O.lfac,
        smem + el * PER, pl, e, first_el, act);
    __syncthreads();  // LDS is reused by the next item
  }
}

// 192 threads: at least three wavefronts per SIMD (the second launch bound),
// every form then fits in 168 VGPRs without spilling (the per-node-velocity
// FMA form would take 174 and hold two); four workgroups of three wavefronts
// fill the CU's twelve.  256 threads (Nq up to 16): two wavefronts per SIMD,
// which the LDS of a Nq = 16 element (78 KB) allows anyway.
template <int NT, int ORDER, bool FMA, bool FACT, bool CV>
__global__ void __launch_bounds__(NT, NT == kEnsNarrow ? 3 : 2) ensemble_kernel(const EnsArgs A) {
  extern __shared__ double smem[];
  int c = 0;
  while (c + 1 < A.ncls && (int)blockIdx.x >= A.cls_block[c + 1]) ++c;
  const int b0 = A.cls_block[c], nb = A.cls_block[c + 1] - b0;
  if ((int)blockIdx.x >= A.cls_block[A.ncls] || nb <= 0) return;  // uniform
  const int lb = blockIdx.x - b0;
  const int first = A.cls_item[c], count = A.cls_item[c + 1] - first, epb = A.cls_epb[c];
  switch (A.cls_nq[c]) {
  case 2: ens_items<NT, 2, ORDER, FMA, FACT, CV>(A, first, count, epb, lb, nb, smem); break;
  case 3: ens_items<NT, 3, ORDER, FMA, FACT, CV>(A, first, count, epb, lb, nb, smem); break;
  case 4: ens_items<NT, 4, ORDER, FMA, FACT, CV>(A, first, count, epb, lb, nb, smem); break;
  case 5: ens_items<NT, 5, ORDER, FMA, FACT, CV>(A, first, count, epb, lb, nb, smem); break;
  case 6: ens_items<NT, 6, ORDER, FMA, FACT, CV>(A, first, count, epb, lb, nb, smem); break;
  case 7: ens_items<NT, 7, ORDER, FMA, FACT, CV>(A, first, count, epb, lb, nb, smem); break;
  case 8: ens_items<NT, 8, ORDER, FMA, FACT, CV>(A, first, count, epb, lb, nb, smem); break;
  case 9: ens_items<NT, 9, ORDER, FMA, FACT, CV>(A, first, count, epb, lb, nb, smem); break;
  case 10: ens_items<NT, 10, ORDER, FMA, FACT, CV>(A, first, count, epb, lb, nb, smem); break;
  case 11: ens_items<NT, 11, ORDER, FMA, FACT, CV>(A, first, count, epb, lb, nb, smem); break;
  case 12: ens_items<NT, 12, ORDER, FMA, FACT, CV>(A, first, count, epb, lb, nb, smem); break;
  case 13: ens_items<NT, 13, ORDER, FMA, FACT, CV>(A, first, count, epb, lb, nb, smem); break;
  default:
    if constexpr (NT >= 256) {
      switch (A.cls_nq[c]) {
      case 14: ens_items<NT, 14, ORDER, FMA, FACT, CV>(A, first, count, epb, lb, nb, smem); break;
      case 15: ens_items<NT, 15, ORDER, FMA, FACT, CV>(A, first, count, epb, lb, nb, smem); break;
      default: ens_items<NT, 16, ORDER, FMA, FACT, CV>(A, first, count, epb, lb, nb, smem); break;
      }
    }
    break;
  }
}

using EnsFn = void (*)(const EnsArgs);

template <int NT, int ORDER, bool FMA, bool FACT>
EnsFn pick_ens_cv(bool cv) {
  return cv ? ensemble_kernel<NT, ORDER, FMA, FACT, true>
            : ensemble_kernel<NT, ORDER, FMA, FACT, false>;
}
template <int NT, int ORDER, bool FMA>
EnsFn pick_ens_fact(bool fact, bool cv) {
  return fact ? pick_ens_cv<NT, ORDER, FMA, true>(cv) : pick_ens_cv<NT, ORDER, FMA, false>(cv);
}
// mixed: the FP32 split contractions (plain products, factored lift)
template <int NT>
EnsFn pick_ens_nt(bool mixed, bool fma, bool fact, bool cv) {
  if (mixed) return pick_ens_cv<NT, kMixed, false, true>(cv);
  return fma ? pick_ens_fact<NT, kOpt1, true>(fact, cv)
             : pick_ens_fact<NT, kOpt1, false>(fact, cv);
}
EnsFn pick_ens(int nt, bool mixed, bool fma, bool fact, bool cv) {
  return nt == kEnsNarrow ? pick_ens_nt<kEnsNarrow>(mixed, fma, fact, cv)
                          : pick_ens_nt<kEnsWide>(mixed, fma, fact, cv);
}

const double kRkA[3] = {0.0, 0.75, 1.0 / 3.0};  // mod_common.f90:29
const double kRkB[3] = {1.0, 0.25, 2.0 / 3.0};  // mod_common.f90:30

}  // namespace

struct dgk_ensemble {
  int n = 0;
  int flags = 0;
  std::vector<EnsInst> inst;   // host copy
  std::vector<int> np;         // nodes per element
  long long total = 0;         // doubles per field pool
  // q and q0 rotate among three buffers as in the single solver: stage 1
  // takes the current buffer as q0, every stage writes the buffer that holds
  // neither its input nor q0
  double *q[3] = {nullptr, nullptr, nullptr};
  double *u = nullptr, *v = nullptr, *w = nullptr;
  int32_t *nbr = nullptr;
  double *diag = nullptr;  // dgk_ensemble_diagnostics scratch (kDiagScratch)
  EnsInst *dinst = nullptr;
  EnsOps *dops = nullptr;
  std::vector<EnsOps> hops;    // host copy
  // stage parameter blocks: [3 cur + c0][stage][instance], rebuilt when dt
  // or a uniform velocity changes
  StageParams *dparams = nullptr;
  int2 *ditems = nullptr;
  int nitems = 0;
  std::vector<int> order;  // instances by increasing order
  int max_nq = 0;
  // classes of equal order, their elements per item and the split of the
  // grid between them
  int ncls = 0;
  int cls_nq[kEnsMaxClass] = {};
  int cls_epb[kEnsMaxClass] = {};
  int cls_item[kEnsMaxClass + 1] = {};
  int cls_block[kEnsMaxClass + 1] = {};
  double cls_weight[kEnsMaxClass] = {};  // DOF, or measured stage time once tuned
  int slots = 0;                         // resident workgroups (multiple of 8)
  bool tuned = false;
  std::vector<double *> opbufs;
  int cur = 0;  // buffer of q
  int c0 = 2;   // buffer of q0 (never cur)
  hipStream_t stream = nullptr;
  EnsFn fn = nullptr;
  int nt = kEnsNarrow;  // threads per workgroup
  size_t lds = 0;
  int grid = 0;
  // one captured step per starting (cur, c0), and the (cur, c0) it ends in
  hipGraphExec_t graph[9] = {};
  int graph_next[9][2] = {};
  bool fields_set = false;
  bool dt_set = false;
};

namespace {

void ens_drop_graphs(dgk_ensemble *s) {
  for (auto &g : s->graph)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
}

// The stage parameter blocks of every instance for each placement (cur, c0)
// of q and q0 and each stage: read q[cur], write the third buffer, q0 is
// q[c0] (stage 1: q[cur] itself).
int ens_upload_params(dgk_ensemble *s) {
  std::vector<StageParams> h((size_t)27 * s->n);
  for (int slot = 0; slot < 27; ++slot) {
    const int cur = slot / 9, c0 = slot / 3 % 3, st = slot % 3;
    if (cur == c0) continue;  // no such placement
    for (int i = 0; i < s->n; ++i) {
      const EnsInst &I = s->inst[i];
      const EnsOps &O = s->hops[I.op];
      StageParams P{};
      P.qin = s->q[cur] + I.off;
      P.qout = s->q[3 - cur - c0] + I.off;
      P.q0 = s->q[st == 0 ? cur : c0] + I.off;
      P.u = s->u + I.off;
      P.v = s->v + I.off;
      P.w = s->w + I.off;
      P.D = O.D;
      P.Lift = O.Lift;
      P.nbr = s->nbr + I.nbr_off;
      for (int f = 0; f < 6; ++f) P.fs[f] = I.fs[f];
      for (int d = 0; d < 3; ++d) {
        P.es[d] = I.es[d];
        P.cvel[d] = I.cvel[d];
      }
      P.Ne = I.ne;
      P.a = kRkA[st];
      P.b = kRkB[st];
      P.dt = I.dt;
      P.mode = st == 0 ? kModeStage1 : kModeStageN;
      h[(size_t)slot * s->n + i] = P;
    }
  }
  HIPOK(hipMemcpy(s->dparams, h.data(), h.size() * sizeof(StageParams), hipMemcpyHostToDevice));
  return DGK_OK;
}

EnsArgs ens_args(const dgk_ensemble *s, int stage) {
  EnsArgs A{};
  A.inst = s->dinst;
  A.ops = s->dops;
  A.items = s->ditems;
  A.params = s->dparams + ((size_t)(9 * s->cur + 3 * s->c0) + stage) * s->n;
  A.ncls = s->ncls;
  for (int c = 0; c < s->ncls; ++c) {
    A.cls_nq[c] = s->cls_nq[c];
    A.cls_epb[c] = s->cls_epb[c];
  }
  for (int c = 0; c <= s->ncls; ++c) {
    A.cls_item[c] = s->cls_item[c];
    A.cls_block[c] = s->cls_block[c];
  }
  return A;
}

int ens_launch_args(dgk_ensemble *s, const EnsArgs &A) {
  const int grid = A.cls_block[A.ncls];
  hipLaunchKernelGGL(s->fn, dim3(grid), dim3(s->nt), s->lds, s->stream, A);
  HIPOK(hipGetLastError());
  return DGK_OK;
}

int ens_launch(dgk_ensemble *s, int stage) {
  const int rc = ens_launch_args(s, ens_args(s, stage));
  if (!rc) {
    const int out = 3 - s->cur - s->c0;
    if (stage == 0) s->c0 = s->cur;
    s->cur = out;
  }
  return rc;
}

// Units of eight workgroups (one per XCD) class c can use: one per eight items.
int ens_unit_cap(const dgk_ensemble *s, int c) {
  return std::max(1, (s->cls_item[c + 1] - s->cls_item[c] + kNumXcd - 1) / kNumXcd);
}

// The grid split as units of eight workgroups per class.
void ens_set_units(dgk_ensemble *s, const int *u) {
  s->cls_block[0] = 0;
  for (int c = 0; c < s->ncls; ++c) s->cls_block[c + 1] = s->cls_block[c] + kNumXcd * u[c];
  s->grid = s->cls_block[s->ncls];
}
void ens_get_units(const dgk_ensemble *s, int *u) {
  for (int c = 0; c < s->ncls; ++c) u[c] = (s->cls_block[c + 1] - s->cls_block[c]) / kNumXcd;
}

// Splits the resident workgroups between the classes in proportion to
// cls_weight (largest remaining weight per unit first), at least one unit
// and at most ens_unit_cap units per class.
void ens_partition(dgk_ensemble *s) {
  const int units = std::max(s->slots / kNumXcd, s->ncls);
  int u[kEnsMaxClass];
  int used = 0;
  for (int c = 0; c < s->ncls; ++c) {
    u[c] = 1;
    ++used;
  }
  while (used < units) {
    int best = -1;
    double bw = -1.0;
    for (int c = 0; c < s->ncls; ++c) {
      if (u[c] >= ens_unit_cap(s, c)) continue;
      const double r = s->cls_weight[c] / u[c];
      if (r > bw) {
        bw = r;
        best = c;
      }
    }
    if (best < 0) break;
    ++u[best];
    ++used;
  }
  ens_set_units(s, u);
}

// The tile shape: nt threads per workgroup and epb[c] elements per item of
// class c (0: as many as fill the workgroup).  Checked before anything
// changes; the work items, the kernel, its LDS and the resident workgroups
// then follow, and the grid is split by the present class weights.  The
// caller has waited for the stream (launches in flight read the items).
int ens_check_shape(const dgk_ensemble *s, int nt, const int *epb) {
  if (nt != kEnsNarrow && nt != kEnsWide)
    return fail(DGK_ERR_ARG, "threads per workgroup %d: %d or %d", nt, kEnsNarrow, kEnsWide);
  if (nt < ens_threads(s->max_nq))
    return fail(DGK_ERR_ARG, "Nq %d (%d lines an element) needs %d threads per workgroup",
                s->max_nq, s->max_nq * s->max_nq, kEnsWide);
  for (int c = 0; epb && c < s->ncls; ++c)
    if (epb[c] < 0 || epb[c] > ens_item_elems(s->cls_nq[c], nt))
      return fail(DGK_ERR_ARG, "class %d (Nq %d): %d elements per item outside 0..%d", c,
                  s->cls_nq[c], epb[c], ens_item_elems(s->cls_nq[c], nt));
  return DGK_OK;
}

int ens_shape(dgk_ensemble *s, int nt, const int *epb) {
  int rc = ens_check_shape(s, nt, epb);
  if (rc) return rc;
  s->nt = nt;
  size_t lds = 0;
  for (int c = 0; c < s->ncls; ++c) {
    const int most = ens_item_elems(s->cls_nq[c], nt);
    s->cls_epb[c] = epb && epb[c] ? epb[c] : most;
    lds = std::max(lds, s->cls_epb[c] * ens_elem_lds(s->cls_nq[c]));
  }
  std::vector<int2> items;
  int c = -1;
  for (int i : s->order) {
    while (c < 0 || s->cls_nq[c] != s->inst[i].nq) s->cls_item[++c] = (int)items.size();
    for (long long e = 0; e < s->inst[i].ne; e += s->cls_epb[c])
      items.push_back(make_int2(i, (int)e));
  }
  s->cls_item[s->ncls] = (int)items.size();
  s->nitems = (int)items.size();
  s->lds = lds;
  HIPOK(hipMemcpy(s->ditems, items.data(), items.size() * sizeof(int2), hipMemcpyHostToDevice));
  const int f = s->flags;
  s->fn = pick_ens(nt, (f & DGK_FLAG_MIXED) != 0, (f & DGK_FLAG_FMA) != 0,
                   (f & DGK_FLAG_FACTORED_LIFT) != 0, (f & DGK_FLAG_CONST_VELOCITY) != 0);
  int dev = 0, ncu = 0, per_cu = 0;
  if (hipFuncSetAttribute((const void *)s->fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)s->lds) != hipSuccess ||
      hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)s->fn, nt, s->lds) !=
          hipSuccess)
    return fail(DGK_ERR_HIP, "ensemble kernel configuration failed");
  s->slots = std::max(1, per_cu) * ncu / kNumXcd * kNumXcd;
  ens_drop_graphs(s);
  ens_partition(s);
  return DGK_OK;
}

// Class weights before any measurement: the DOF of each class.
void ens_dof(const dgk_ensemble *s, double *w) {
  for (int c = 0; c < s->ncls; ++c) w[c] = 0.0;
  for (int i = 0; i < s->n; ++i)
    for (int c = 0; c < s->ncls; ++c)
      if (s->cls_nq[c] == s->inst[i].nq) w[c] += (double)s->inst[i].ne * s->np[i];
}
void ens_dof_weights(dgk_ensemble *s) { ens_dof(s, s->cls_weight); }

template <class T>
int ens_alloc(T **p, size_t n) {
  *p = nullptr;
  if (!n) return DGK_OK;
  if (hipMalloc((void **)p, n * sizeof(T)) != hipSuccess) {
    *p = nullptr;
    return fail(DGK_ERR_NOMEM, "hipMalloc of %zu bytes failed", n * sizeof(T));
  }
  return DGK_OK;
}

}  // namespace

extern "C" {

int dgk_ensemble_destroy(dgk_ensemble *s) {
  if (!s) return DGK_OK;
  ens_drop_graphs(s);
  void *bufs[] = {s->diag, s->q[0], s->q[1], s->q[2], s->u, s->v, s->w, s->nbr, s->dinst, s->dops, s->ditems,
                  s->dparams};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  for (double *b : s->opbufs)
    if (b) (void)hipFree(b);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
  return DGK_OK;
}

int dgk_ensemble_create(dgk_ensemble **out, int n, const int32_t *p, const int32_t *nex,
                        const int32_t *ney, const int32_t *nez, const double *dom,
                        const char *operator_dir, int type_id, int flags) {
  clear_error();
  if (!out || n < 1 || !p || !nex || !ney || !nez || !operator_dir)
    return fail(DGK_ERR_ARG, "bad argument");
  *out = nullptr;
  if (type_id != DGK_TYPE_OPT1 && type_id != DGK_TYPE_HIP)
    return fail(DGK_ERR_OPTTYPE, "the ensemble evaluates in OPT1's summation order only");
  if (flags & ~(DGK_FLAG_FMA | DGK_FLAG_FACTORED_LIFT | DGK_FLAG_CONST_VELOCITY | DGK_FLAG_MIXED))
    return fail(DGK_ERR_ARG, "flags 0x%x not supported by the ensemble", flags);
  if ((flags & DGK_FLAG_MIXED) && (flags & DGK_FLAG_FMA))
    return fail(DGK_ERR_ARG, "DGK_FLAG_MIXED runs plain FP32 products (no DGK_FLAG_FMA)");
  if (flags & DGK_FLAG_MIXED) flags |= DGK_FLAG_FACTORED_LIFT;  // the split form lifts in 1D
  for (int i = 0; i < n; ++i) {
    int32_t sz[7];
    int rc = dgk_mesh_sizes(nex[i], ney[i], nez[i], p[i], sz);
    if (rc) return rc;
    for (int d = 0; dom && d < 3; ++d)
      if (!(dom[3 * i + d] > 0.0) || dom[3 * i + d] > 1e300)
        return fail(DGK_ERR_ARG, "instance %d: domain length %g is not positive and finite", i,
                    dom[3 * i + d]);
    if (!ens_supported(p[i] + 1))
      return fail(DGK_ERR_ORDER, "instance %d: PolyOrder %d outside 1..15", i, p[i]);
  }
  dgk_ensemble *s = new dgk_ensemble;
  s->n = n;
  s->flags = flags;
  s->inst.resize(n);
  s->np.resize(n);
  int rc = DGK_OK;
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
    delete s;
    return fail(DGK_ERR_HIP, "hipStreamCreate failed");
  }
  // operators, one slot per distinct order
  std::map<int, int> slot;
  std::vector<EnsOps> &hops = s->hops;
  for (int i = 0; i < n && !rc; ++i) {
    if (slot.count(p[i])) continue;
    const int Nq = p[i] + 1, Np = Nq * Nq * Nq;
    std::vector<double> x1D(Nq), D(Nq * Nq), L(6 * (size_t)Np);
    if ((rc = dgk_read_operator(operator_dir, p[i], x1D.data(), D.data(), L.data()))) break;
    std::vector<double> Drow(Nq * Nq), LiftP(lift_delta_doubles(Nq)), lf(2 * Nq);
    for (int k = 0; k < Nq; ++k)
      for (int l = 0; l < Nq; ++l) Drow[k * Nq + l] = D[k + l * Nq];
    if (!pack_lift_deltas(Nq, L.data(), LiftP.data()) && !(flags & DGK_FLAG_FACTORED_LIFT)) {
      rc = fail(DGK_ERR_ARG, "PolyOrder %d: the exact lift needs a Lift whose values along each "
                "face's normal index lie within 255 units in the last place of one another; "
                "add DGK_FLAG_FACTORED_LIFT", p[i]);
      break;
    }
    for (int t = 0; t < Nq; ++t) {
      lf[t] = L[t + (size_t)Np * 3];
      lf[Nq + t] = L[t + (size_t)Np * 1];
    }
    // FP32 split copies and line-fit nodes for DGK_FLAG_MIXED, in buffers of
    // doubles' room (the device reads the split parts as floats)
    std::vector<double> d32(Nq * Nq + 3 * Nq);
    split_tables(Nq, Drow.data(), d32.data());
    EnsOps o{};
    const std::vector<double> *src[6] = {&D, &L, &Drow, &LiftP, &lf, &d32};
    const double **dst[6] = {&o.D, &o.Lift, &o.Drow, &o.LiftP, &o.lfac, &o.Drow32};
    for (int b = 0; b < 6 && !rc; ++b) {
      double *d = nullptr;
      if ((rc = ens_alloc(&d, src[b]->size()))) break;
      s->opbufs.push_back(d);
      if (hipMemcpy(d, src[b]->data(), src[b]->size() * sizeof(double), hipMemcpyHostToDevice) !=
          hipSuccess)
        rc = fail(DGK_ERR_HIP, "operator upload failed");
      *dst[b] = d;
    }
    slot[p[i]] = (int)hops.size();
    hops.push_back(o);
  }
  // instances: geometry constants, offsets, neighbour tables
  std::vector<int32_t> nbr;
  long long off = 0;
  for (int i = 0; i < n && !rc; ++i) {
    const int Nq = p[i] + 1;
    EnsInst &I = s->inst[i];
    I.nq = Nq;
    I.op = slot[p[i]];
    I.ne = (long long)nex[i] * ney[i] * nez[i];
    I.off = off;
    I.nbr_off = (long long)nbr.size();
    const double L3[3] = {dom ? dom[3 * i] : 1.0, dom ? dom[3 * i + 1] : 1.0,
                          dom ? dom[3 * i + 2] : 1.0};
    const double dl[3] = {L3[0] / nex[i], L3[1] / ney[i], L3[2] / nez[i]};
    for (int d = 0; d < 3; ++d) I.es[d] = 2.0 / dl[d];                      // mod_mesh.f90:147-149
    for (int f = 0; f < 6; ++f) I.fs[f] = 2.0 / dl[kFace[f].axis];          // mod_mesh.f90:366-394
    I.cvel[0] = I.cvel[1] = I.cvel[2] = 0.0;
    I.dt = 0.0;
    s->np[i] = Nq * Nq * Nq;
    off += I.ne * s->np[i];
    for (int k = 0; k < nez[i]; ++k)
      for (int j = 0; j < ney[i]; ++j)
        for (int ii = 0; ii < nex[i]; ++ii)
          for (int f = 0; f < 6; ++f) {
            bool wrapped;
            nbr.push_back((int32_t)neighbour_elem(f, ii, j, k, nex[i], ney[i], nez[i], &wrapped));
          }
  }
  s->total = off;
  // classes of equal order; work items (instance, first element) sorted by
  // order, then instance, then element, built by ens_shape
  s->order.resize(n);
  for (int i = 0; i < n; ++i) s->order[i] = i;
  std::stable_sort(s->order.begin(), s->order.end(),
                   [&](int a, int b) { return s->inst[a].nq < s->inst[b].nq; });
  long long nelem = 0;
  for (int i : s->order) {
    const int nq = s->inst[i].nq;
    if (!s->ncls || s->cls_nq[s->ncls - 1] != nq) s->cls_nq[s->ncls++] = nq;
    s->max_nq = std::max(s->max_nq, nq);
    nelem += s->inst[i].ne;
  }
  ens_dof_weights(s);
  if (!rc) rc = ens_alloc(&s->q[0], off);
  if (!rc) rc = ens_alloc(&s->q[1], off);
  if (!rc) rc = ens_alloc(&s->q[2], off);
  if (!rc) rc = ens_alloc(&s->u, off);
  if (!rc) rc = ens_alloc(&s->v, off);
  if (!rc) rc = ens_alloc(&s->w, off);
  if (!rc) rc = ens_alloc(&s->nbr, nbr.size());
  if (!rc) rc = ens_alloc(&s->dinst, (size_t)n);
  if (!rc) rc = ens_alloc(&s->dops, hops.size());
  if (!rc) rc = ens_alloc(&s->ditems, (size_t)nelem);  // one element per item at most
  if (!rc) rc = ens_alloc(&s->dparams, (size_t)27 * n);
  if (!rc && (hipMemcpy(s->nbr, nbr.data(), nbr.size() * sizeof(int32_t), hipMemcpyHostToDevice) ||
              hipMemcpy(s->dops, hops.data(), hops.size() * sizeof(EnsOps), hipMemcpyHostToDevice) ||
              hipMemcpy(s->dinst, s->inst.data(), n * sizeof(EnsInst), hipMemcpyHostToDevice)))
    rc = fail(DGK_ERR_HIP, "ensemble table upload failed");
  if (!rc && hipMemset(s->q[2], 0, off * sizeof(double)) != hipSuccess)
    rc = fail(DGK_ERR_HIP, "hipMemset failed");
  if (!rc) rc = ens_upload_params(s);
  if (!rc) rc = ens_shape(s, ens_threads(s->max_nq), nullptr);
  if (rc) {
    dgk_ensemble_destroy(s);
    return rc;
  }
  *out = s;
  return DGK_OK;
}

int dgk_ensemble_set_fields(dgk_ensemble *s, int i, const double *q, const double *u,
                            const double *v, const double *w) {
  if (!s || i < 0 || i >= s->n || !q || !u || !v || !w) return fail(DGK_ERR_ARG, "bad argument");
  HIPOK(hipStreamSynchronize(s->stream));  // steps in flight still read the fields
  EnsInst &I = s->inst[i];
  const size_t cnt = (size_t)I.ne * s->np[i];
  if (s->flags & DGK_FLAG_CONST_VELOCITY) {
    const double *vel[3] = {u, v, w};
    for (int d = 0; d < 3; ++d) {
      size_t t = 0;
      if (cnt > 1 && find_not_equal(vel[d] + 1, cnt - 1, vel[d][0], &t))
        return fail(DGK_ERR_ARG, "instance %d: velocity component %d is not uniform", i, d + 1);
      I.cvel[d] = vel[d][0];
    }
    HIPOK(hipMemcpy(s->dinst + i, &I, sizeof(EnsInst), hipMemcpyHostToDevice));
    int rc = ens_upload_params(s);
    if (rc) return rc;
  }
  double *dst[4] = {s->q[s->cur] + I.off, s->u + I.off, s->v + I.off, s->w + I.off};
  const double *src[4] = {q, u, v, w};
  for (int b = 0; b < 4; ++b) {
    const int rc = copy_h2d(dst[b], src[b], cnt * sizeof(double), s->stream);
    if (rc) return rc;
  }
  s->fields_set = true;
  return DGK_OK;
}

int dgk_ensemble_get_field(dgk_ensemble *s, int i, double *q) {
  if (!s || i < 0 || i >= s->n || !q) return fail(DGK_ERR_ARG, "bad argument");
  HIPOK(hipStreamSynchronize(s->stream));
  const EnsInst &I = s->inst[i];
  return copy_d2h(q, s->q[s->cur] + I.off, (size_t)I.ne * s->np[i] * sizeof(double), s->stream);
}

int dgk_ensemble_set_dt(dgk_ensemble *s, const double *dt) {
  if (!s || !dt) return fail(DGK_ERR_ARG, "bad argument");
  HIPOK(hipStreamSynchronize(s->stream));
  for (int i = 0; i < s->n; ++i)
    if (!(dt[i] >= 0.0) || dt[i] > 1e300)
      return fail(DGK_ERR_ARG, "instance %d: dt %g is not a finite non-negative number", i, dt[i]);
  for (int i = 0; i < s->n; ++i) s->inst[i].dt = dt[i];
  s->dt_set = true;
  HIPOK(hipMemcpy(s->dinst, s->inst.data(), s->n * sizeof(EnsInst), hipMemcpyHostToDevice));
  return ens_upload_params(s);
}

int dgk_ensemble_step(dgk_ensemble *s) {
  if (!s) return fail(DGK_ERR_ARG, "null ensemble");
  if (!s->fields_set) return fail(DGK_ERR_STATE, "no fields set");
  if (!s->dt_set) return fail(DGK_ERR_STATE, "no time steps set (dgk_ensemble_set_dt)");
  const int key = 3 * s->cur + s->c0, cur0 = s->cur, c00 = s->c0;
  if (!s->graph[key]) {
    hipGraph_t g;
    HIPOK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    int rc = DGK_OK;
    for (int st = 0; st < 3 && !rc; ++st) rc = ens_launch(s, st);
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
  }
  HIPOK(hipGraphLaunch(s->graph[key], s->stream));
  s->cur = s->graph_next[key][0];
  s->c0 = s->graph_next[key][1];
  return DGK_OK;
}

int dgk_ensemble_run(dgk_ensemble *s, int nstep) {
  for (int it = 0; it < nstep; ++it) {
    int rc = dgk_ensemble_step(s);
    if (rc) return rc;
  }
  HIPOK(hipStreamSynchronize(s->stream));
  return DGK_OK;
}

int dgk_ensemble_time_stages(dgk_ensemble *s, int nstage, double *ms_per_stage) {
  if (!s || nstage < 1 || !ms_per_stage) return fail(DGK_ERR_ARG, "bad argument");
  if (nstage % 3)
    return fail(DGK_ERR_ARG, "nstage %d is not a whole number of SSP-RK3 steps (3 stages each)",
                nstage);
  if (!s->fields_set) return fail(DGK_ERR_STATE, "no fields set");
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = DGK_OK;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess ||
      hipEventRecord(e0, s->stream) != hipSuccess)
    rc = fail(DGK_ERR_HIP, "ensemble time_stages: event setup failed");
  for (int it = 0; it < nstage && !rc; ++it) rc = ens_launch(s, it % 3);
  float ms = 0.f;
  if (!rc && (hipEventRecord(e1, s->stream) != hipSuccess ||
              hipEventSynchronize(e1) != hipSuccess ||
              hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
    rc = fail(DGK_ERR_HIP, "ensemble time_stages: event timing failed");
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (rc) return rc;
  *ms_per_stage = ms / nstage;
  return DGK_OK;
}

}  // extern "C"

namespace {

// Average time of nrep launches of A (after one untimed launch).
int ens_time_args(dgk_ensemble *s, const EnsArgs &A, int nrep, hipEvent_t e0, hipEvent_t e1,
                  double *ms) {
  int rc = ens_launch_args(s, A);
  if (rc) return rc;
  HIPOK(hipEventRecord(e0, s->stream));
  for (int r = 0; r < nrep && !rc; ++r) rc = ens_launch_args(s, A);
  if (rc) return rc;
  float f = 0.f;
  HIPOK(hipEventRecord(e1, s->stream));
  HIPOK(hipEventSynchronize(e1));
  HIPOK(hipEventElapsedTime(&f, e0, e1));
  *ms = (double)f / nrep;
  return DGK_OK;
}

// Stage 1 of class c alone on as much of the grid as it can use: reads q and
// writes the buffer that holds neither q nor q0, so the state is left as it
// was.  *wg_ms: the time times the workgroups it had, the share of the grid
// the class needs for an equal finish time.
int ens_time_class(dgk_ensemble *s, int c, int nrep, hipEvent_t e0, hipEvent_t e1, double *ms,
                   double *wg_ms) {
  EnsArgs A = ens_args(s, 0);
  const int u = std::min(std::max(s->slots / kNumXcd, 1), ens_unit_cap(s, c));
  A.ncls = 1;
  A.cls_nq[0] = s->cls_nq[c];
  A.cls_epb[0] = s->cls_epb[c];
  A.cls_item[0] = s->cls_item[c];
  A.cls_item[1] = s->cls_item[c + 1];
  A.cls_block[0] = 0;
  A.cls_block[1] = kNumXcd * u;
  int rc = ens_time_args(s, A, nrep, e0, e1, ms);
  *wg_ms = *ms * u;
  return rc;
}

// The tile-shape search of dgk_ensemble_tune.  For each workgroup width
// that the largest order allows, each class in turn tries its candidate
// elements per item (all that fill the workgroup, about half, one) with the
// other classes at their choice so far, and keeps the one with the least
// workgroup-time (time alone times the workgroups it had).  The grid is then
// split between the classes in proportion to those workgroup-times, to their
// DOF, and to the geometric mean of the two, the whole stage is timed for
// each, and the fastest width and split are kept.  Last, one unit of eight
// workgroups at a time moves between the largest class and each other one
// while the whole stage gets faster (classes timed alone misjudge what they
// take from each other when they share the chip).
int ens_tune(dgk_ensemble *s, int nrep, double *ms_per_class, hipEvent_t e0, hipEvent_t e1) {
  int widths[2], nw = 0;
  if (s->max_nq <= 13) widths[nw++] = kEnsNarrow;
  widths[nw++] = kEnsWide;
  double best_stage = 0.0;
  int best_nt = 0, best_epb[kEnsMaxClass] = {}, best_u[kEnsMaxClass] = {};
  double best_ms[kEnsMaxClass] = {};
  double dof[kEnsMaxClass], dsum = 0.0;
  ens_dof(s, dof);
  for (int c = 0; c < s->ncls; ++c) dsum += dof[c];
  for (int k = 0; k < nw; ++k) {
    const int nt = widths[k];
    int epb[kEnsMaxClass];
    double w[kEnsMaxClass], ms[kEnsMaxClass];
    for (int c = 0; c < s->ncls; ++c) epb[c] = ens_item_elems(s->cls_nq[c], nt);
    for (int c = 0; c < s->ncls; ++c) {
      const int most = epb[c];
      const int cand[3] = {most, (most + 1) / 2, 1};
      double bw = 0.0, bms = 0.0;
      int be = most;
      for (int j = 0; j < 3; ++j) {
        if (j && (cand[j] == cand[j - 1] || cand[j] == most)) continue;
        epb[c] = cand[j];
        double t = 0.0, wt = 0.0;
        int rc = ens_shape(s, nt, epb);
        if (!rc) rc = ens_time_class(s, c, nrep, e0, e1, &t, &wt);
        if (rc) return rc;
        if (j == 0 || wt < bw) {
          bw = wt;
          bms = t;
          be = cand[j];
        }
      }
      epb[c] = be;
      w[c] = bw;
      ms[c] = bms;
    }
    int rc = ens_shape(s, nt, epb);
    if (rc) return rc;
    double wsum = 0.0;
    for (int c = 0; c < s->ncls; ++c) wsum += w[c];
    for (int split = 0; split < 3; ++split) {
      for (int c = 0; c < s->ncls; ++c) {
        const double a = w[c] / wsum, b = dof[c] / dsum;
        s->cls_weight[c] = split == 0 ? a : split == 1 ? b : std::sqrt(a * b);
      }
      ens_partition(s);
      double stage = 0.0;
      if ((rc = ens_time_args(s, ens_args(s, 0), nrep, e0, e1, &stage))) return rc;
      if ((k == 0 && split == 0) || stage < best_stage) {
        best_stage = stage;
        best_nt = nt;
        ens_get_units(s, best_u);
        for (int c = 0; c < s->ncls; ++c) {
          best_epb[c] = epb[c];
          best_ms[c] = ms[c];
        }
      }
    }
  }
  int rc = ens_shape(s, best_nt, best_epb);
  if (rc) return rc;
  ens_set_units(s, best_u);
  int big = 0;
  for (int c = 1; c < s->ncls; ++c)
    if (best_u[c] > best_u[big]) big = c;
  for (int c = 0; c < s->ncls; ++c) {
    if (c == big) continue;
    for (int dir = 1; dir >= -1; dir -= 2) {
      for (int step = 0; step < 4; ++step) {  // while it pays, up to four units
        int u[kEnsMaxClass];
        for (int d = 0; d < s->ncls; ++d) u[d] = best_u[d];
        u[c] += dir;
        u[big] -= dir;
        if (u[c] < 1 || u[big] < 1 || u[c] > ens_unit_cap(s, c) || u[big] > ens_unit_cap(s, big))
          break;
        ens_set_units(s, u);
        double stage = 0.0;
        if ((rc = ens_time_args(s, ens_args(s, 0), nrep, e0, e1, &stage))) return rc;
        if (!(stage < best_stage * 0.998)) break;
        best_stage = stage;
        for (int d = 0; d < s->ncls; ++d) best_u[d] = u[d];
      }
    }
  }
  // the split kept as weights that ens_partition maps back to these units
  for (int c = 0; c < s->ncls; ++c) {
    s->cls_weight[c] = best_u[c];
    if (ms_per_class) ms_per_class[c] = best_ms[c];
  }
  ens_set_units(s, best_u);
  return DGK_OK;
}

}  // namespace

extern "C" {

int dgk_ensemble_tune(dgk_ensemble *s, int nrep, double *ms_per_class) {
  if (!s || nrep < 1) return fail(DGK_ERR_ARG, "bad argument");
  if (!s->fields_set) return fail(DGK_ERR_STATE, "no fields set");
  HIPOK(hipStreamSynchronize(s->stream));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = DGK_OK;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    rc = fail(DGK_ERR_HIP, "ensemble tune: hipEventCreate failed");
  if (!rc) rc = ens_tune(s, nrep, ms_per_class, e0, e1);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (rc) {  // back to the untuned shape
    ens_dof_weights(s);
    s->tuned = false;
    (void)ens_shape(s, ens_threads(s->max_nq), nullptr);
    return rc;
  }
  s->tuned = true;
  return DGK_OK;
}

int dgk_ensemble_set_shape(dgk_ensemble *s, int threads, const int32_t *elems_per_item) {
  if (!s) return fail(DGK_ERR_ARG, "null ensemble");
  const int nt = threads ? threads : ens_threads(s->max_nq);
  int epb[kEnsMaxClass];
  for (int c = 0; c < s->ncls; ++c) epb[c] = elems_per_item ? elems_per_item[c] : 0;
  int rc = ens_check_shape(s, nt, epb);
  if (rc) return rc;
  HIPOK(hipStreamSynchronize(s->stream));
  ens_dof_weights(s);
  s->tuned = false;
  return ens_shape(s, nt, epb);
}

int dgk_ensemble_diagnostics(dgk_ensemble *s, int i, const double *w1D, const double *ref,
                             double *out) {
  if (!s || i < 0 || i >= s->n || !w1D || !out) return fail(DGK_ERR_ARG, "bad argument");
  if (!s->fields_set) return fail(DGK_ERR_STATE, "no fields set");
  if (!s->diag) {
    int rc = ens_alloc(&s->diag, (size_t)kDiagScratch);
    if (rc) return rc;
  }
  const EnsInst &I = s->inst[i];
  return device_diagnostics(s->q[s->cur] + I.off, ref, I.ne * s->np[i], I.nq, I.es, nullptr,
                            w1D, s->diag, s->stream, out);
}

int dgk_ensemble_query(dgk_ensemble *s, int what, int64_t *value) {
  if (!s || !value) return fail(DGK_ERR_ARG, "bad argument");
  switch (what) {
  case 0: *value = s->total; break;          // DOF over all instances
  case 1: *value = s->nitems; break;         // work items per stage
  case 2: *value = s->grid; break;           // workgroups of the persistent launch
  case 3: *value = (int64_t)s->lds; break;   // LDS bytes per workgroup
  case 4: *value = s->ncls; break;           // classes of equal order
  case 5: *value = s->tuned; break;          // grid split from measured times
  case 6: *value = s->nt; break;             // threads per workgroup
  default:
    if (what >= 16 && what < 16 + s->ncls) {  // Nq of class c = what-16
      *value = s->cls_nq[what - 16];
    } else if (what >= 48 && what < 48 + s->ncls) {  // elements per item of class c
      *value = s->cls_epb[what - 48];
    } else if (what >= 32 && what < 32 + s->ncls) {  // workgroups of class c
      *value = s->cls_block[what - 32 + 1] - s->cls_block[what - 32];
    } else {
      return fail(DGK_ERR_ARG, "unknown query %d", what);
    }
  }
  return DGK_OK;
}

}  // extern "C"
