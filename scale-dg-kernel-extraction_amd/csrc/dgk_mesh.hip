// Device-side build of the periodic structured mesh tables
// (mod_mesh.f90:74-193, 332-397) straight into device memory: VMapM, VMapP,
// normal_fn, Escale, Fscale, pos_en and halo_src_map in the reference's
// numbering, the same values as the host build (dgk_mesh_setup).
//
// The reference numbers halo slots in one serial pass (elements in order,
// faces -y,+x,+y,-x,-z,+z, face nodes in order; mod_mesh.f90:171-190).  Here
// every face node computes its slot in closed form: the number of wrapped
// faces of all earlier elements plus the wrapped faces before it in its own
// element, times Nfp, plus its face-node index.
#include "dgk_internal.hpp"

#include <hip/hip_runtime.h>

using namespace dgk;

#define HIPOK(expr)                                                            \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess)                                                      \
      return fail(DGK_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

namespace {

struct MeshArgs {
  int NeX, NeY, NeZ, Nq, Np, Nfp, NfpTot;
  long long Ne;
  double dl[3];
  double x1D[16];  // LGL nodes (Nq <= 16), passed by value
  int32_t *VMapM, *VMapP, *halo;
  double *nrm, *escale, *fscale, *pos;
};

// Face f: normal axis, normal sign, opposite face (mod_mesh.f90:269-276).
__device__ __forceinline__ int f_axis(int f) { return f < 4 ? ((f & 1) ? 0 : 1) : 2; }
__device__ __forceinline__ int f_sign(int f) { return (f == 1 || f == 2 || f == 5) ? 1 : -1; }
__device__ __forceinline__ int f_opp(int f) { return f ^ (f < 4 ? 2 : 1); }

// Local node of face f at tangential (a, b) (Fmask, mod_mesh.f90:278-324).
__device__ __forceinline__ int f_node(int Nq, int f, int a, int b) {
  int i, j, k;
  switch (f) {
  case 0: i = a; j = 0; k = b; break;
  case 1: i = Nq - 1; j = a; k = b; break;
  case 2: i = a; j = Nq - 1; k = b; break;
  case 3: i = 0; j = a; k = b; break;
  case 4: i = a; j = b; k = 0; break;
  default: i = a; j = b; k = Nq - 1; break;
  }
  return i + j * Nq + k * Nq * Nq;
}

// Wrapped faces of an element along one axis of length n at coordinate c,
// and the same count summed over the coordinates before c.
__device__ __forceinline__ int wrap1(int c, int n) { return (c == 0) + (c == n - 1); }
__device__ __forceinline__ long long wrap_before(int c) { return c > 0 ? 1 : 0; }

// Wrapped faces of all elements before (i,j,k) in the reference's order
// (i fastest).  Every full line of an axis holds two wrapped faces of it.
__device__ __forceinline__ long long wrapped_before(const MeshArgs &A, int i, int j, int k) {
  const long long X = A.NeX, Y = A.NeY;
  // full z-planes k' < k
  long long w = (long long)k * (2 * Y + 2 * X) + X * Y * wrap_before(k);
  // full rows j' < j of plane k
  w += (long long)j * (2 + X * wrap1(k, A.NeZ)) + X * wrap_before(j);
  // elements i' < i of row j
  w += wrap_before(i) + (long long)i * (wrap1(j, A.NeY) + wrap1(k, A.NeZ));
  return w;
}

__global__ void __launch_bounds__(256) mesh_face_kernel(const MeshArgs A) {
  const long long total = A.Ne * A.NfpTot;
  const long long NfNe = total;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    const long long e = t / A.NfpTot;
    const int r = (int)(t - e * A.NfpTot);
    const int f = r / A.Nfp, fp = r - f * A.Nfp;
    const int i = (int)(e % A.NeX), j = (int)((e / A.NeX) % A.NeY),
              k = (int)(e / ((long long)A.NeX * A.NeY));
    const int ax = f_axis(f), sg = f_sign(f);
    int c[3] = {i, j, k};
    const int n[3] = {A.NeX, A.NeY, A.NeZ};
    c[ax] += sg;
    bool wrapped = false;
    if (c[ax] < 0) { c[ax] = n[ax] - 1; wrapped = true; }
    if (c[ax] >= n[ax]) { c[ax] = 0; wrapped = true; }
    const long long eP = c[0] + (long long)c[1] * A.NeX + (long long)c[2] * A.NeX * A.NeY;
    const int a = fp % A.Nq, b = fp / A.Nq;
    if (A.VMapM) A.VMapM[t] = (int32_t)(f_node(A.Nq, f, a, b) + 1 + e * A.Np);
    const int32_t src = (int32_t)(f_node(A.Nq, f_opp(f), a, b) + 1 + eP * A.Np);
    if (wrapped) {
      // wrapped faces before face f within the element, in face order
      int before = 0;
      for (int g = 0; g < f; ++g) {
        const int gx = f_axis(g), gs = f_sign(g);
        const int cc = gx == 0 ? i : (gx == 1 ? j : k);
        before += gs < 0 ? (cc == 0) : (cc == n[gx] - 1);
      }
      const long long slot = (wrapped_before(A, i, j, k) + before) * A.Nfp + fp;  // 0-based
      if (A.VMapP) A.VMapP[t] = (int32_t)(A.Np * A.Ne + slot + 1);
      if (A.halo) A.halo[slot] = src;
    } else if (A.VMapP) {
      A.VMapP[t] = src;
    }
    if (A.nrm)
      for (int d = 0; d < 3; ++d) A.nrm[t + d * NfNe] = d == ax ? (double)sg : 0.0;
    if (A.fscale) A.fscale[t] = 2.0 / A.dl[ax];
  }
}

__global__ void __launch_bounds__(256) mesh_node_kernel(const MeshArgs A) {
  const long long NpNe = A.Ne * A.Np;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < NpNe;
       t += (long long)gridDim.x * blockDim.x) {
    const long long e = t / A.Np;
    const int nd = (int)(t - e * A.Np);
    const int ie[3] = {(int)(e % A.NeX), (int)((e / A.NeX) % A.NeY),
                       (int)(e / ((long long)A.NeX * A.NeY))};
    const int id[3] = {nd % A.Nq, (nd / A.Nq) % A.Nq, nd / (A.Nq * A.Nq)};
    for (int d = 0; d < 3; ++d) {
      if (A.pos) {
        const double c = ((double)(ie[d] + 1) - 0.5) * A.dl[d];  // mod_mesh.f90:344-347
        A.pos[t + d * NpNe] = c + 0.5 * A.dl[d] * A.x1D[id[d]];
      }
      if (A.escale) A.escale[t + d * NpNe] = 2.0 / A.dl[d];  // mod_mesh.f90:147-149
    }
  }
}

}  // namespace

extern "C" int dgk_mesh_setup_device(int NeX, int NeY, int NeZ, int p, double domx,
                                     double domy, double domz, const double *x1D,
                                     int32_t *VMapM, int32_t *VMapP, double *normal_fn,
                                     double *Escale, double *Fscale, double *pos_en,
                                     int32_t *halo_src_map, void *stream) {
  clear_error();
  int32_t sz[7];
  int rc = dgk_mesh_sizes(NeX, NeY, NeZ, p, sz);
  if (rc) return rc;
  if (pos_en && !x1D) return fail(DGK_ERR_ARG, "pos_en requires x1D");
  MeshArgs A{};
  A.NeX = NeX;
  A.NeY = NeY;
  A.NeZ = NeZ;
  A.Nq = sz[0];
  A.Np = sz[1];
  A.Nfp = sz[2];
  A.NfpTot = sz[3];
  A.Ne = sz[4];
  A.dl[0] = domx / (double)NeX;
  A.dl[1] = domy / (double)NeY;
  A.dl[2] = domz / (double)NeZ;
  A.VMapM = VMapM;
  A.VMapP = VMapP;
  A.halo = halo_src_map;
  A.nrm = normal_fn;
  A.escale = Escale;
  A.fscale = Fscale;
  A.pos = pos_en;
  hipStream_t st = (hipStream_t)stream;
  if (pos_en)
    for (int i = 0; i < A.Nq; ++i) A.x1D[i] = x1D[i];
  const long long nf = A.Ne * A.NfpTot, nn = A.Ne * A.Np;
  const unsigned gf = (unsigned)std::min<long long>((nf + 255) / 256, 65536);
  const unsigned gn = (unsigned)std::min<long long>((nn + 255) / 256, 65536);
  if (VMapM || VMapP || normal_fn || Fscale || halo_src_map)
    hipLaunchKernelGGL(mesh_face_kernel, dim3(gf), dim3(256), 0, st, A);
  if (pos_en || Escale) hipLaunchKernelGGL(mesh_node_kernel, dim3(gn), dim3(256), 0, st, A);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return fail(DGK_ERR_HIP, "mesh kernels: %s", hipGetErrorString(e));
  return DGK_OK;
}
