// Internal declarations shared by the host C++ and the HIP translation units.
#pragma once

#include "../../include/dgk.h"

namespace dgk {

// Face f (0-based: 0 -y, 1 +x, 2 +y, 3 -x, 4 -z, 5 +z): normal axis, normal
// sign and opposite face (mod_mesh.f90:269-276, 366-394, 420-480).
struct FaceInfo {
  int axis;
  int sign;
  int opp;
};
extern const FaceInfo kFace[6];

int face_node_host(int Nq, int f, int a, int b);
long long neighbour_elem(int f, int i, int j, int k, int NeX, int NeY, int NeZ,
                         bool *wrapped);

// Records a message for dgk_last_error() and returns code.
int fail(int code, const char *fmt, ...);
void clear_error();

// FP32 split copies of the row-major operator drow (Nq x Nq) and of the
// factored-lift vectors lf = [l_-, l_+] (2 Nq) for DGK_FLAG_MIXED: x = hi +
// lo, hi = fl32(x), lo = fl32(x - hi).  Row k of d32 (2 Nq floats) holds
// D_hi(k,.) then D_lo(k,.); l32 (4 Nq floats) holds lf's hi parts then its lo
// parts.
void split_tables(int Nq, const double *drow, const double *lf, float *d32, float *l32);

// The device reductions of dgk_diagnostics over the n = Np Ne owned values
// of q (elements of Nq^3 nodes; uniform Escale es, or the per-node escale
// (Np,Ne,3) if not null), against the device field ref (or null): out[8] as
// dgk_diagnostics documents.  scratch: kDiagScratch device doubles; stream:
// the hipStream_t to run on.  Waits for the result.
constexpr long long kDiagScratch = 8 * 1024 + 8;  // 1024 block partials, the result
int device_diagnostics(const double *q, const double *ref, long long n, int Nq,
                       const double es[3], const double *escale, const double *w1D,
                       double *scratch, void *stream, double *out);

}  // namespace dgk
