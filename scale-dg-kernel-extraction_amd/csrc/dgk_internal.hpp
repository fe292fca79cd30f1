// Internal declarations shared by the host C++ and the HIP translation units.
#pragma once

#include "../../include/dgk.h"

#include <cstddef>

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

// The dense lift Lift(i,j,k,f) (the reference's (Np,6) layout) packed for
// the pencil kernels without loss: each value is the double whose bit
// pattern is its group's base plus an unsigned byte, the group being face f's
// values with one normal index n (j for faces 0 and 2, i for 1 and 3, k for
// 4 and 5), so that a lane reads 8 bytes per node instead of 6 doubles.
// out (lift_delta_doubles(Nq) doubles): the 6 Nq bases (as doubles, group
// f Nq + n: the group's smallest bit pattern), then one 8-byte word per
// node, its byte f the delta of Lift(i,j,k,f) (bytes 6, 7 zero): for node k
// of line pl = i + Nq j word (k - k%2) Nq^2 + 2 pl + k%2, so that the words
// of nodes 2r and 2r+1 of a line are one 16-byte load, except for the last
// node of an odd Nq, word (Nq-1) Nq^2 + pl.  False (out unspecified) when a
// group mixes signs, holds a non-finite value or spans more than 255 units
// in the last place.
constexpr size_t lift_delta_doubles(int Nq) { return 6 * (size_t)Nq + (size_t)Nq * Nq * Nq; }
bool pack_lift_deltas(int Nq, const double *Lift, double *out);

// Line-fit tables of the matrix-core mixed form (Nq = 16): out[8 n + f] =
// l_f(x_n), the Lagrange basis on the nodes f of kMfmaFitNodes at LGL node
// n, and out[128 + 8 n + f] = sum_l D(n,l) l_f(x_l), from the row-major
// operator drow.  kMfmaFitDoubles doubles.
constexpr int kMfmaFitNodes[8] = {0, 2, 4, 6, 9, 11, 13, 15};  // fit_node in dgk_mfma.hip
constexpr int kMfmaFitDoubles = 256;
void mfma_fit_tables(const double *drow, double *out);

// The device reductions of dgk_diagnostics over the n = Np Ne owned values
// of q (elements of Nq^3 nodes; uniform Escale es, or the per-node escale
// (Np,Ne,3) if not null), against the device field ref (or null): out[8] as
// dgk_diagnostics documents.  scratch: kDiagScratch device doubles; stream:
// the hipStream_t to run on.  Waits for the result.
constexpr long long kDiagScratch = 8 * 1024 + 8;  // 1024 block partials, the result
int device_diagnostics(const double *q, const double *ref, long long n, int Nq,
                       const double es[3], const double *escale, const double *w1D,
                       double *scratch, void *stream, double *out);

// Synchronous copies of `bytes` between host memory (pageable) and device
// memory on `stream` (a hipStream_t; its earlier work is ordered before the
// copy): large copies go through two pinned staging buffers, filled and
// emptied by several host threads while the DMA of the other buffer runs.
int copy_h2d(void *dst, const void *src, size_t bytes, void *stream);
int copy_d2h(void *dst, const void *src, size_t bytes, void *stream);
// 1 and *where = the first index i < n with a[i] != v (NaN never equals),
// else 0 and *where = n; the scan is split between the same host threads.
int find_not_equal(const double *a, size_t n, double v, size_t *where);

}  // namespace dgk
