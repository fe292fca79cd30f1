// Host-side producers of the DG element and mesh data: operator file reader,
// selector, and the periodic structured hexahedral mesh in the reference's
// numbering.  Pure C++ (no device code).
//
// Reference numbering reproduced here (mod_mesh.f90):
//   element id  ke = i + (j-1)NeX + (k-1)NeX NeY            (:486-493)
//   node id     n  = i + (j-1)Nq + (k-1)Nq^2                (:267)
//   faces       1 -y, 2 +x, 3 +y, 4 -x, 5 -z, 6 +z          (:269-276)
//   face nodes  Fmask(fp,f), fp running over the two tangential indices,
//               first index fastest                          (:278-324)
//   VMapM = Fmask(fp,f) + (ke-1)Np; interior VMapP = Fmask(fp,opp) +
//   (keP-1)Np; periodic boundary faces point into a halo buffer numbered in
//   (k,j,i,f,fp) loop order                                  (:153-190)
#include "dgk_internal.hpp"

#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace dgk {

static thread_local std::string g_last_error;

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

void clear_error() { g_last_error.clear(); }

const FaceInfo kFace[6] = {
    // axis, sign, opposite
    {1, -1, 2},  // 1: -y
    {0, +1, 3},  // 2: +x
    {1, +1, 0},  // 3: +y
    {0, -1, 1},  // 4: -x
    {2, -1, 5},  // 5: -z
    {2, +1, 4},  // 6: +z
};

// Local (0-based) node of face f (0-based) at tangential position (a, b).
int face_node_host(int Nq, int f, int a, int b) {
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

// Neighbour across face f of element (i,j,k) (0-based grid coordinates) on a
// periodic NeX x NeY x NeZ grid; *wrapped is set when the face lies on the
// domain boundary.
long long neighbour_elem(int f, int i, int j, int k, int NeX, int NeY, int NeZ,
                         bool *wrapped) {
  int c[3] = {i, j, k};
  const int n[3] = {NeX, NeY, NeZ};
  const int ax = kFace[f].axis;
  c[ax] += kFace[f].sign;
  *wrapped = false;
  if (c[ax] < 0) { c[ax] = n[ax] - 1; *wrapped = true; }
  if (c[ax] >= n[ax]) { c[ax] = 0; *wrapped = true; }
  return c[0] + (long long)c[1] * NeX + (long long)c[2] * NeX * NeY;
}

// Legendre-Gauss-Lobatto nodes on [-1,1] (ascending): x_0 = -1, x_{Nq-1} = 1
// and the roots of P'_{Nq-1} between them, by Newton's iteration on
// x P_N(x) - P_{N-1}(x), which vanishes at all of them, from the
// Chebyshev-Gauss-Lobatto points (the method of G. von Winckel's lglnodes).
void lgl_nodes(int Nq, double *x) {
  const int N = Nq - 1;
  for (int j = 0; j < Nq; ++j) {
    double xj = -std::cos(M_PI * j / N), prev = 2.0;
    for (int it = 0; it < 100 && std::fabs(xj - prev) > 1e-16; ++it) {
      double p0 = 1.0, p1 = xj;  // P_{k-1}, P_k
      for (int k = 2; k <= N; ++k) {
        const double p2 = ((2 * k - 1) * xj * p1 - (k - 1) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      prev = xj;
      if (j == 0 || j == N) break;
      xj = xj - (xj * p1 - p0) / (Nq * p1);
    }
    x[j] = xj;
  }
  x[0] = -1.0;
  x[N] = 1.0;
}

// Layout (see dgk_internal.hpp): the line fit of DGK_FLAG_MIXED takes out
// the polynomial of degree deg = min(3, Nq-1) through the nodes 0, a, b and
// Nq-1 (a = Nq/3, b = 2Nq/3) in the Newton basis 1, t, p2 = t(t-1),
// p3 = t(t-1)(t-t_a) of t = (x+1)/2; the remainder vanishes at those nodes,
// so their slots in the split rows are free and carry the basis
// derivatives d p2/dx and d p3/dx at the row's node (as two dwords each).
void split_tables(int Nq, const double *drow, double *d32) {
  const int a = Nq / 3, b = 2 * Nq / 3;
  std::vector<double> x(Nq), t(Nq);
  lgl_nodes(Nq, x.data());
  for (int l = 0; l < Nq; ++l) t[l] = 0.5 * (x[l] + 1.0);
  t[0] = 0.0;
  t[Nq - 1] = 1.0;
  auto p2 = [&](double s) { return s * (s - 1.0); };
  auto p3 = [&](double s) { return s * (s - 1.0) * (s - t[a]); };
  // derivatives with respect to x (dt/dx = 1/2)
  auto q2 = [&](double s) { return 0.5 * (2.0 * s - 1.0); };
  auto q3 = [&](double s) { return 0.5 * ((2.0 * s - 1.0) * (s - t[a]) + s * (s - 1.0)); };
  std::vector<float> rows(2 * (size_t)Nq * Nq);
  for (int k = 0; k < Nq; ++k) {
    float *hi = rows.data() + (size_t)k * 2 * Nq, *lo = hi + Nq;
    for (int l = 0; l < Nq; ++l) {
      const double v = drow[(size_t)k * Nq + l];
      hi[l] = (float)v;
      lo[l] = (float)(v - (double)hi[l]);
    }
    uint32_t w2[2], w3[2];
    const double d2 = q2(t[k]), d3 = q3(t[k]);
    std::memcpy(w2, &d2, sizeof d2);
    std::memcpy(w3, &d3, sizeof d3);
    std::memcpy(&hi[0], &w2[0], 4);
    std::memcpy(&hi[Nq - 1], &w2[1], 4);
    std::memcpy(&lo[0], &w3[0], 4);
    std::memcpy(&lo[Nq - 1], &w3[1], 4);
  }
  std::memcpy(d32, rows.data(), rows.size() * sizeof(float));
  double *T = d32 + (size_t)Nq * Nq, *P2 = T + Nq, *P3 = P2 + Nq;
  for (int l = 0; l < Nq; ++l) {
    T[l] = t[l];
    P2[l] = p2(t[l]);
    P3[l] = p3(t[l]);
  }
  // the end slots of T (t = 0 and 1 there, never read) hold 1/p2(t_a) and
  // 1/p3(t_b), the Newton divisors
  T[0] = Nq >= 3 ? 1.0 / p2(t[a]) : 0.0;
  T[Nq - 1] = Nq >= 4 ? 1.0 / p3(t[b]) : 0.0;
}


// Face f's normal index at node (i,j,k) of an element: j for -y and +y
// (f = 0, 2), i for +x and -x (f = 1, 3), k for -z and +z (f = 4, 5).
static int lift_normal_index(int f, int i, int j, int k) {
  return f == 0 || f == 2 ? j : f == 1 || f == 3 ? i : k;
}

bool pack_lift_deltas(int Nq, const double *Lift, double *out) {
  const int Nfp = Nq * Nq, Np = Nfp * Nq;
  auto bits = [](double x) {
    int64_t b;
    std::memcpy(&b, &x, sizeof b);
    return b;
  };
  // the groups (f, n): every value of face f's column with normal index n
  std::vector<int64_t> lo(6 * Nq, INT64_MAX), hi(6 * Nq, INT64_MIN);
  std::vector<int> neg(6 * Nq, 0), cnt(6 * Nq, 0);
  for (int f = 0; f < 6; ++f)
    for (int k = 0; k < Nq; ++k)
      for (int j = 0; j < Nq; ++j)
        for (int i = 0; i < Nq; ++i) {
          const double x = Lift[i + Nq * j + Nfp * k + (size_t)Np * f];
          const int g = f * Nq + lift_normal_index(f, i, j, k);
          const int64_t b = bits(x);
          lo[g] = std::min(lo[g], b);
          hi[g] = std::max(hi[g], b);
          neg[g] += std::signbit(x) ? 1 : 0;
          ++cnt[g];
          if (!std::isfinite(x)) return false;
        }
  std::vector<int64_t> base(6 * Nq);
  for (int g = 0; g < 6 * Nq; ++g) {
    if (neg[g] != 0 && neg[g] != cnt[g]) return false;  // one sign per group
    if (hi[g] - lo[g] > 255) return false;
    base[g] = lo[g];
    std::memcpy(&out[g], &base[g], sizeof(double));
  }
  auto *bytes = reinterpret_cast<uint8_t *>(out + 6 * Nq);
  std::memset(bytes, 0, (size_t)Np * 8);
  for (int pl = 0; pl < Nfp; ++pl)
    for (int k = 0; k < Nq; ++k) {
      // the words of nodes 2r and 2r+1 of a line side by side; an odd Nq's
      // last node alone
      const size_t word = (Nq % 2 == 1 && k == Nq - 1)
                              ? (size_t)k * Nfp + pl
                              : (size_t)(k & ~1) * Nfp + 2 * (size_t)pl + (k & 1);
      for (int f = 0; f < 6; ++f) {
        const int i = pl % Nq, j = pl / Nq;
        const int64_t d =
            bits(Lift[pl + Nfp * k + (size_t)Np * f]) - base[f * Nq + lift_normal_index(f, i, j, k)];
        bytes[word * 8 + f] = (uint8_t)d;
      }
    }
  return true;
}


void mfma_fit_tables(const double *drow, double *out) {
  constexpr int Nq = 16;
  double x[Nq];
  lgl_nodes(Nq, x);
  double *lb = out, *gd = out + 128;
  for (int n = 0; n < Nq; ++n)
    for (int f = 0; f < 8; ++f) {
      double v = 1.0;
      const double xf = x[kMfmaFitNodes[f]];
      for (int h = 0; h < 8; ++h)
        if (h != f) v *= (x[n] - x[kMfmaFitNodes[h]]) / (xf - x[kMfmaFitNodes[h]]);
      lb[8 * n + f] = v;
    }
  for (int n = 0; n < Nq; ++n)
    for (int f = 0; f < 8; ++f) {
      double s = 0.0;
      for (int l = 0; l < Nq; ++l) s += drow[n * Nq + l] * lb[8 * l + f];
      gd[8 * n + f] = s;
    }
}

}  // namespace dgk

using namespace dgk;

extern "C" {

const char *dgk_last_error(void) { return g_last_error.c_str(); }

int dgk_version(void) { return 10000; }

int dgk_optr_kernel_setup(const char *evaltype, int *typeid_out) {
  if (!evaltype || !typeid_out) return fail(DGK_ERR_ARG, "null argument");
  std::string s(evaltype);
  while (!s.empty() && (s.back() == ' ' || s.back() == '\0')) s.pop_back();
  if (s == "GENERAL") *typeid_out = DGK_TYPE_GENERAL;
  else if (s == "OPT1") *typeid_out = DGK_TYPE_OPT1;
  else if (s == "HIP") *typeid_out = DGK_TYPE_HIP;
  else return fail(DGK_ERR_OPTTYPE, "Unsupported dgoptr_evaltype: %s", s.c_str());
  return DGK_OK;
}

int dgk_read_operator(const char *dir, int p, double *x1D, double *D1D,
                      double *Lift_mat) {
  if (p < 1 || p > 15) return fail(DGK_ERR_ORDER, "PolyOrder %d outside 1..15", p);
  if (!dir || !x1D || !D1D || !Lift_mat) return fail(DGK_ERR_ARG, "null argument");
  std::string path = std::string(dir) + "/p" + std::to_string(p) + ".dat";
  FILE *fp = std::fopen(path.c_str(), "r");
  if (!fp) return fail(DGK_ERR_IO, "cannot open %s", path.c_str());
  const long Nq = p + 1;
  const long counts[3] = {Nq, Nq * Nq, 6 * Nq * Nq * Nq};
  double *dst[3] = {x1D, D1D, Lift_mat};
  for (int a = 0; a < 3; ++a)
    for (long i = 0; i < counts[a]; ++i)
      if (std::fscanf(fp, "%lf", &dst[a][i]) != 1) {
        std::fclose(fp);
        return fail(DGK_ERR_IO, "%s: expected %ld values in record %d", path.c_str(),
                    counts[a], a + 1);
      }
  std::fclose(fp);
  return DGK_OK;
}

int dgk_mesh_sizes(int NeX, int NeY, int NeZ, int p, int32_t *sizes) {
  if (p < 1 || p > 15) return fail(DGK_ERR_ORDER, "PolyOrder %d outside 1..15", p);
  if (NeX < 1 || NeY < 1 || NeZ < 1 || !sizes) return fail(DGK_ERR_ARG, "bad mesh size");
  const long long Nq = p + 1, Np = Nq * Nq * Nq, Nfp = Nq * Nq;
  const long long Ne = (long long)NeX * NeY * NeZ;
  const long long Nhalo = 2LL * ((long long)NeY * NeZ + (long long)NeX * NeZ + (long long)NeX * NeY);
  if (Np * (Ne + Nhalo) >= (1LL << 31))
    return fail(DGK_ERR_ARG, "Np*NeA = %lld exceeds the int32 index range of VMapM/VMapP",
                Np * (Ne + Nhalo));
  sizes[0] = (int32_t)Nq;
  sizes[1] = (int32_t)Np;
  sizes[2] = (int32_t)Nfp;
  sizes[3] = (int32_t)(6 * Nfp);
  sizes[4] = (int32_t)Ne;
  sizes[5] = (int32_t)(Ne + Nhalo);
  sizes[6] = (int32_t)(Nfp * Nhalo);
  return DGK_OK;
}

int dgk_mesh_setup(int NeX, int NeY, int NeZ, int p, double domx, double domy,
                   double domz, const double *x1D, int32_t *VMapM,
                   int32_t *VMapP, double *normal_fn, double *Escale,
                   double *Fscale, double *pos_en, int32_t *halo_src_map) {
  int32_t sz[7];
  int rc = dgk_mesh_sizes(NeX, NeY, NeZ, p, sz);
  if (rc) return rc;
  if (pos_en && !x1D) return fail(DGK_ERR_ARG, "pos_en requires x1D");
  const int Nq = sz[0], Np = sz[1], Nfp = sz[2], NfpTot = sz[3];
  const long long Ne = sz[4];
  const double dl[3] = {domx / (double)NeX, domy / (double)NeY, domz / (double)NeZ};
  const long long NfNe = (long long)NfpTot * Ne, NpNe = (long long)Np * Ne;
  std::vector<int> fnode(6 * Nfp);
  for (int f = 0; f < 6; ++f)
    for (int fp = 0; fp < Nfp; ++fp) fnode[f * Nfp + fp] = face_node_host(Nq, f, fp % Nq, fp / Nq);

  long long ibuf = 0;
  for (int k = 0; k < NeZ; ++k)
    for (int j = 0; j < NeY; ++j)
      for (int i = 0; i < NeX; ++i) {
        const long long e = i + (long long)j * NeX + (long long)k * NeX * NeY;
        const double c[3] = {((double)(i + 1) - 0.5) * dl[0], ((double)(j + 1) - 0.5) * dl[1],
                             ((double)(k + 1) - 0.5) * dl[2]};
        if (pos_en)
          for (int n = 0; n < Np; ++n) {
            const int id[3] = {n % Nq, (n / Nq) % Nq, n / (Nq * Nq)};
            for (int d = 0; d < 3; ++d)
              pos_en[n + e * Np + d * NpNe] = c[d] + 0.5 * dl[d] * x1D[id[d]];
          }
        if (Escale)
          for (int d = 0; d < 3; ++d)
            for (int n = 0; n < Np; ++n) Escale[n + e * Np + d * NpNe] = 2.0 / dl[d];
        for (int f = 0; f < 6; ++f) {
          bool wrapped;
          const long long eP = neighbour_elem(f, i, j, k, NeX, NeY, NeZ, &wrapped);
          const int opp = kFace[f].opp;
          for (int fp = 0; fp < Nfp; ++fp) {
            const long long t = f * Nfp + fp + e * NfpTot;
            if (VMapM) VMapM[t] = (int32_t)(fnode[f * Nfp + fp] + 1 + e * Np);
            const int32_t src = (int32_t)(fnode[opp * Nfp + fp] + 1 + eP * Np);
            if (wrapped) {
              ++ibuf;
              if (VMapP) VMapP[t] = (int32_t)(Np * Ne + ibuf);
              if (halo_src_map) halo_src_map[ibuf - 1] = src;
            } else if (VMapP) {
              VMapP[t] = src;
            }
            if (normal_fn)
              for (int d = 0; d < 3; ++d)
                normal_fn[t + d * NfNe] = (d == kFace[f].axis) ? (double)kFace[f].sign : 0.0;
            if (Fscale) Fscale[t] = 2.0 / dl[kFace[f].axis];
          }
        }
      }
  return DGK_OK;
}

}  // extern "C"
