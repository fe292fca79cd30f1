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

// Tables of DGK_FLAG_MIXED.  d32 (Nq^2 + 3 Nq doubles' room): row k of the
// row-major operator drow (Nq x Nq) as 2 Nq floats, D_hi(k,.) then D_lo(k,.)
// (x = hi + lo, hi = fl32(x), lo = fl32(x - hi)), rows filling the first Nq^2
// doubles, with the basis derivatives of the line fit packed into the slots
// of the fit nodes; then three rows of Nq doubles over the LGL nodes x_l:
// t_l = (x_l + 1)/2 (its end slots holding the fit's two Newton divisors),
// p2(t_l) and p3(t_l) (dgk_host.cpp gives the basis).
void split_tables(int Nq, const double *drow, double *d32);
// Legendre-Gauss-Lobatto nodes x[Nq] on [-1,1], ascending.
void lgl_nodes(int Nq, double *x);

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
