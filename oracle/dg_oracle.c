/*
 * dg_oracle.c -- CPU restatement of the SCALE-DG 3D DG advection tendency and
 * SSP-RK3 step loop (Masahiro-Ren/scale-dg-kernel-extraction).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle: only tests/, the smoke
 * check in __graft_entry__.py and the cpu_baseline leg of bench.py may load
 * it, and only as the checker or the timed CPU baseline.  The product path
 * (the HIP library) never calls it.
 *
 * Every routine repeats the reference's floating-point operations in the
 * reference's order (no reassociation; build with -ffp-contract=off) so that
 * its results are bit-identical to a non-FMA build of the reference.  Array
 * layouts are the reference's Fortran column-major layouts, flattened:
 *   q(Np,NeA)                q[n + ke*Np]
 *   VMapM/VMapP(NfpTot,Ne)   1-based flat indices into q(Np*NeA)
 *   normal_fn(NfpTot,Ne,3)   [t + ke*NfpTot + d*NfpTot*Ne]
 *   Escale(Np,Ne,3)          [n + ke*Np + d*Np*Ne]
 *   Fscale(NfpTot,Ne)        [t + ke*NfpTot]
 *   pos_en(Np,Ne,3)          [n + ke*Np + d*Np*Ne]
 *   D1D(Nq,Nq)               D(i,l) at [i + l*Nq]
 *   Lift_mat(Nq,Nq,Nq,6)     [i + j*Nq + k*Nq^2 + f*Np]
 *   Fmask(Nfp,6)             1-based node indices, [fp + f*Nfp]
 *
 * Pinning: bit-exact against the flang-compiled reference (golden fixtures
 * under tests/golden/), plus closed-form checks (LGL nodes, D*x^m, exact mass
 * lift, conservation, RK amplification polynomial) in tests/.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_GENERAL 1 /* mod_dg_optr_kernel.f90:39 */
#define ORC_OPT1 2    /* mod_dg_optr_kernel.f90:42 */

/* SSP-RK3 Shu-Osher coefficients, mod_common.f90:28-30 */
static const double rk_a[3] = {0.0, 0.75, 1.0 / 3.0};
static const double rk_b[3] = {1.0, 0.25, 2.0 / 3.0};

/* ------------------------------------------------------------------------ */
/* Operator data: mod_mesh.f90:253-263 (list-directed read of x1D, D1D, Lift) */
int orc_read_operator(const char *dir, int p, double *x1D, double *D1D,
                      double *Lift) {
  char path[1024];
  snprintf(path, sizeof path, "%s/p%d.dat", dir, p);
  FILE *fp = fopen(path, "r");
  if (!fp) return 1;
  int Nq = p + 1;
  long n1 = Nq, n2 = (long)Nq * Nq, n3 = 6L * Nq * Nq * Nq;
  for (long i = 0; i < n1; ++i)
    if (fscanf(fp, "%lf", &x1D[i]) != 1) { fclose(fp); return 2; }
  for (long i = 0; i < n2; ++i)
    if (fscanf(fp, "%lf", &D1D[i]) != 1) { fclose(fp); return 2; }
  for (long i = 0; i < n3; ++i)
    if (fscanf(fp, "%lf", &Lift[i]) != 1) { fclose(fp); return 2; }
  fclose(fp);
  return 0;
}

/* Face node tables, mod_mesh.f90:278-324.  Faces: 1 -y, 2 +x, 3 +y, 4 -x,
 * 5 -z, 6 +z.  Node n = i + (j-1)Nq + (k-1)Nq^2 (1-based). */
void orc_fmask(int Nq, int *Fmask) {
  int Nfp = Nq * Nq, fp;
  int Nq2 = Nq * Nq;
  fp = 0;
  for (int k = 1; k <= Nq; ++k)
    for (int i = 1; i <= Nq; ++i) Fmask[fp++ + 0 * Nfp] = i + (k - 1) * Nq2;
  fp = 0;
  for (int k = 1; k <= Nq; ++k)
    for (int j = 1; j <= Nq; ++j)
      Fmask[fp++ + 1 * Nfp] = Nq + (j - 1) * Nq + (k - 1) * Nq2;
  fp = 0;
  for (int k = 1; k <= Nq; ++k)
    for (int i = 1; i <= Nq; ++i)
      Fmask[fp++ + 2 * Nfp] = i + (Nq - 1) * Nq + (k - 1) * Nq2;
  fp = 0;
  for (int k = 1; k <= Nq; ++k)
    for (int j = 1; j <= Nq; ++j)
      Fmask[fp++ + 3 * Nfp] = 1 + (j - 1) * Nq + (k - 1) * Nq2;
  fp = 0;
  for (int j = 1; j <= Nq; ++j)
    for (int i = 1; i <= Nq; ++i) Fmask[fp++ + 4 * Nfp] = i + (j - 1) * Nq;
  fp = 0;
  for (int j = 1; j <= Nq; ++j)
    for (int i = 1; i <= Nq; ++i)
      Fmask[fp++ + 5 * Nfp] = i + (j - 1) * Nq + (Nq - 1) * Nq2;
}

/* elem_id, mod_mesh.f90:486-493 */
static int elem_id(int i, int j, int k, int NeX, int NeY) {
  return i + (j - 1) * NeX + (k - 1) * NeX * NeY;
}

/* eval_neighbor_element, mod_mesh.f90:407-483 */
static void neighbor(int i, int j, int k, int f, int NeX, int NeY, int NeZ,
                     int *keP, int *opp, int *is_bnd) {
  *is_bnd = 0;
  switch (f) {
  case 1:
    *opp = 3;
    if (j > 1) *keP = elem_id(i, j - 1, k, NeX, NeY);
    else { *keP = elem_id(i, NeY, k, NeX, NeY); *is_bnd = 1; }
    break;
  case 2:
    *opp = 4;
    if (i < NeX) *keP = elem_id(i + 1, j, k, NeX, NeY);
    else { *keP = elem_id(1, j, k, NeX, NeY); *is_bnd = 1; }
    break;
  case 3:
    *opp = 1;
    if (j < NeY) *keP = elem_id(i, j + 1, k, NeX, NeY);
    else { *keP = elem_id(i, 1, k, NeX, NeY); *is_bnd = 1; }
    break;
  case 4:
    *opp = 2;
    if (i > 1) *keP = elem_id(i - 1, j, k, NeX, NeY);
    else { *keP = elem_id(NeX, j, k, NeX, NeY); *is_bnd = 1; }
    break;
  case 5:
    *opp = 6;
    if (k > 1) *keP = elem_id(i, j, k - 1, NeX, NeY);
    else { *keP = elem_id(i, j, NeZ, NeX, NeY); *is_bnd = 1; }
    break;
  default:
    *opp = 5;
    if (k < NeZ) *keP = elem_id(i, j, k + 1, NeX, NeY);
    else { *keP = elem_id(i, j, 1, NeX, NeY); *is_bnd = 1; }
    break;
  }
}

/* Sizes of mesh_setup, mod_mesh.f90:96-100 and 238-241. */
void orc_mesh_sizes(int NeX, int NeY, int NeZ, int p, int *sizes) {
  int Nq = p + 1, Nfp = Nq * Nq;
  int Ne = NeX * NeY * NeZ;
  int Nhalo = 2 * (NeY * NeZ + NeX * NeZ + NeX * NeY);
  sizes[0] = Nq;
  sizes[1] = Nq * Nq * Nq;   /* Np */
  sizes[2] = Nfp;            /* Nfp */
  sizes[3] = 6 * Nfp;        /* NfpTot */
  sizes[4] = Ne;
  sizes[5] = Ne + Nhalo;     /* NeA */
  sizes[6] = Nfp * Nhalo;    /* NhaloNode */
}

/* mesh_setup + set_element_position + set_face_geometry,
 * mod_mesh.f90:74-193, 332-397.  Outputs must be preallocated with the sizes
 * from orc_mesh_sizes. */
void orc_mesh_setup(int NeX, int NeY, int NeZ, int p, double domx, double domy,
                    double domz, const double *x1D, int *VMapM, int *VMapP,
                    double *normal_fn, double *Escale, double *Fscale,
                    double *pos_en, int *halo_src_map) {
  int Nq = p + 1, Np = Nq * Nq * Nq, Nfp = Nq * Nq, NfpTot = 6 * Nfp;
  int Ne = NeX * NeY * NeZ;
  int *Fmask = (int *)malloc(sizeof(int) * Nfp * 6);
  orc_fmask(Nq, Fmask);
  double dx = domx / (double)NeX;
  double dy = domy / (double)NeY;
  double dz = domz / (double)NeZ;
  long NfNe = (long)NfpTot * Ne, NpNe = (long)Np * Ne;

  for (int k = 1; k <= NeZ; ++k)
    for (int j = 1; j <= NeY; ++j)
      for (int i = 1; i <= NeX; ++i) {
        int ke = elem_id(i, j, k, NeX, NeY);
        long e0 = ke - 1;
        double xc = ((double)i - 0.5) * dx;
        double yc = ((double)j - 0.5) * dy;
        double zc = ((double)k - 0.5) * dz;
        for (int kk = 0; kk < Nq; ++kk)
          for (int jj = 0; jj < Nq; ++jj)
            for (int ii = 0; ii < Nq; ++ii) {
              long n = ii + jj * Nq + kk * Nq * Nq;
              pos_en[n + e0 * Np + 0 * NpNe] = xc + 0.5 * dx * x1D[ii];
              pos_en[n + e0 * Np + 1 * NpNe] = yc + 0.5 * dy * x1D[jj];
              pos_en[n + e0 * Np + 2 * NpNe] = zc + 0.5 * dz * x1D[kk];
            }
        for (int n = 0; n < Np; ++n) {
          Escale[n + e0 * Np + 0 * NpNe] = 2.0 / dx;
          Escale[n + e0 * Np + 1 * NpNe] = 2.0 / dy;
          Escale[n + e0 * Np + 2 * NpNe] = 2.0 / dz;
        }
        /* set_face_geometry */
        for (int d = 0; d < 3; ++d)
          for (int t = 0; t < NfpTot; ++t) normal_fn[t + e0 * NfpTot + d * NfNe] = 0.0;
        static const int nd[6] = {1, 0, 1, 0, 2, 2};
        static const double ns[6] = {-1.0, 1.0, 1.0, -1.0, -1.0, 1.0};
        double fs[6] = {2.0 / dy, 2.0 / dx, 2.0 / dy, 2.0 / dx, 2.0 / dz, 2.0 / dz};
        for (int f = 0; f < 6; ++f)
          for (int fp = 0; fp < Nfp; ++fp) {
            long t = f * Nfp + fp;
            normal_fn[t + e0 * NfpTot + nd[f] * NfNe] = ns[f];
            Fscale[t + e0 * NfpTot] = fs[f];
          }
        for (int f = 1; f <= 6; ++f) {
          int keP, opp, bnd;
          neighbor(i, j, k, f, NeX, NeY, NeZ, &keP, &opp, &bnd);
          for (int fp = 1; fp <= Nfp; ++fp)
            VMapM[(f - 1) * Nfp + fp - 1 + e0 * NfpTot] =
                Fmask[fp - 1 + (f - 1) * Nfp] + (ke - 1) * Np;
          if (!bnd)
            for (int fp = 1; fp <= Nfp; ++fp)
              VMapP[(f - 1) * Nfp + fp - 1 + e0 * NfpTot] =
                  Fmask[fp - 1 + (opp - 1) * Nfp] + (keP - 1) * Np;
        }
      }
  /* serial halo numbering, mod_mesh.f90:171-190 */
  int ibuf = 0;
  for (int k = 1; k <= NeZ; ++k)
    for (int j = 1; j <= NeY; ++j)
      for (int i = 1; i <= NeX; ++i) {
        int ke = elem_id(i, j, k, NeX, NeY);
        for (int f = 1; f <= 6; ++f) {
          int keP, opp, bnd;
          neighbor(i, j, k, f, NeX, NeY, NeZ, &keP, &opp, &bnd);
          for (int fp = 1; fp <= Nfp; ++fp)
            if (bnd) {
              ibuf += 1;
              VMapP[(f - 1) * Nfp + fp - 1 + (long)(ke - 1) * NfpTot] = Np * Ne + ibuf;
              halo_src_map[ibuf - 1] = Fmask[fp - 1 + (opp - 1) * Nfp] + (keP - 1) * Np;
            }
        }
      }
  free(Fmask);
}

/* update_halo, mod_mesh.f90:209-221 (serial in the reference) */
void orc_update_halo(double *f, int Np, int Ne, int NhaloNode,
                     const int *halo_src_map) {
  long base = (long)Np * Ne;
  for (int ib = 0; ib < NhaloNode; ++ib)
    f[base + ib] = f[halo_src_map[ib] - 1];
}

/* cal_elembnd_flux, mod_advect3d_eq.f90:89-142 */
void orc_elembnd_flux(double *flux, const double *q, const double *u,
                      const double *v, const double *w, const int *VMapM,
                      const int *VMapP, const double *nrm, const double *Fscale,
                      int Np, int NfpTot, int Ne) {
  long NfNe = (long)NfpTot * Ne;
  (void)Np;
#pragma omp parallel for schedule(static)
  for (long ke = 0; ke < Ne; ++ke)
    for (int t = 0; t < NfpTot; ++t) {
      long o = t + ke * NfpTot;
      long iM = VMapM[o] - 1, iP = VMapP[o] - 1;
      double n1 = nrm[o], n2 = nrm[o + NfNe], n3 = nrm[o + 2 * NfNe];
      double qM = q[iM], qP = q[iP];
      double VelM = u[iM] * n1 + v[iM] * n2 + w[iM] * n3;
      double VelP = u[iP] * n1 + v[iP] * n2 + w[iP] * n3;
      double alpha = 0.5 * fabs(VelP + VelM);
      flux[o] = 0.5 * Fscale[o] * (qP * VelP - qM * VelM - alpha * (qP - qM));
    }
}

/* Split point of OPT1 for p > 7: Ruby ((p+1)/3.0).round,
 * mod_dg_optr_kernel_opt1.F90.erb:43 */
int orc_opt1_split(int Nq) { return (int)floor(Nq / 3.0 + 0.5); }

/* GENERAL sweeps, mod_dg_optr_kernel.f90:197-258: out = 0; out += a*b for
 * l = 1..Nq.  OPT1 sweeps, mod_dg_optr_kernel_opt1.F90.erb:228-344: a
 * left-to-right sum for p <= 7, three partial sums for p > 7. */
static double sum_terms(const double *a, long sa, const double *b, long sb,
                        int Nq, int typeid, int x_order) {
  /* x_order: the product is written Mat(i,l)*in(l) (x sweep) rather than
   * in(..l..)*Mat_tr(l,.) (y,z sweeps); IEEE multiplication commutes, so the
   * flag documents the reference form only. */
  (void)x_order;
  if (typeid == ORC_GENERAL) {
    double s = 0.0;
    for (int l = 0; l < Nq; ++l) s = s + a[l * sa] * b[l * sb];
    return s;
  }
  if (Nq <= 8) {
    double s = a[0] * b[0];
    for (int l = 1; l < Nq; ++l) s = s + a[l * sa] * b[l * sb];
    return s;
  }
  int r = orc_opt1_split(Nq);
  double t1 = a[0] * b[0];
  for (int l = 1; l < r; ++l) t1 = t1 + a[l * sa] * b[l * sb];
  double t2 = a[r * sa] * b[r * sb];
  for (int l = r + 1; l < 2 * r; ++l) t2 = t2 + a[l * sa] * b[l * sb];
  double t3 = a[2 * r * sa] * b[2 * r * sb];
  for (int l = 2 * r + 1; l < Nq; ++l) t3 = t3 + a[l * sa] * b[l * sb];
  return t1 + t2 + t3;
}

/* tensorprod_divlike_dirXYZ for one element (mod_dg_optr_kernel.f90:62-136) */
void orc_divlike(int Nq, int typeid, const double *Mat, const double *Mat_tr,
                 const double *in_x, const double *in_y, const double *in_z,
                 double *out_x, double *out_y, double *out_z) {
  int Nq2 = Nq * Nq;
  for (int k = 0; k < Nq; ++k)
    for (int j = 0; j < Nq; ++j)
      for (int i = 0; i < Nq; ++i) {
        long n = i + j * Nq + (long)k * Nq2;
        /* x: Mat(i,l) * in_x(l,jk) */
        out_x[n] = sum_terms(&Mat[i], Nq, &in_x[j * Nq + (long)k * Nq2], 1, Nq, typeid, 1);
        /* y: in_y(i,l,k) * Mat_tr(l,j) */
        out_y[n] = sum_terms(&in_y[i + (long)k * Nq2], Nq, &Mat_tr[j * Nq], 1, Nq, typeid, 0);
        /* z: in_z(i,j,l) * Mat_tr(l,k) */
        out_z[n] = sum_terms(&in_z[i + j * Nq], Nq2, &Mat_tr[k * Nq], 1, Nq, typeid, 0);
      }
}

/* Left-to-right sum a(1)*b(1) + a(2)*b(2) + ... + a(Nq)*b(Nq) with no split
 * at any p: the directional kernels dirX/dirY/dirZ and gradlike_dirXYZ of the
 * OPT1 module (mod_dg_optr_kernel_opt1.F90.erb:63-66, 92-95, 121-124,
 * 184-206). */
static double sum_seq(const double *a, long sa, const double *b, long sb, int Nq) {
  double s = a[0] * b[0];
  for (int l = 1; l < Nq; ++l) s = s + a[l * sa] * b[l * sb];
  return s;
}

/* element_operation_kernel_matvec_dir{X,Y,Z}_P<p> for one element
 * (mod_dg_optr_kernel_opt1.F90.erb:51-129): dir 0 out(i,jk) = Mat(i,l)
 * in(l,jk); dir 1 out(i,j,k) = in(i,l,k) Mat(l,j); dir 2 out(i,j,k) =
 * in(i,j,l) Mat(l,k).  Mat is Mat_x for x and the transposed Mat_y / Mat_z
 * for y and z, as the reference's callers pass them. */
void orc_matvec_dir(int Nq, int dir, const double *Mat, const double *in, double *out) {
  int Nq2 = Nq * Nq;
  for (int k = 0; k < Nq; ++k)
    for (int j = 0; j < Nq; ++j)
      for (int i = 0; i < Nq; ++i) {
        long n = i + j * Nq + (long)k * Nq2;
        if (dir == 0)
          out[n] = sum_seq(&Mat[i], Nq, &in[j * Nq + (long)k * Nq2], 1, Nq);
        else if (dir == 1)
          out[n] = sum_seq(&in[i + (long)k * Nq2], Nq, &Mat[j * Nq], 1, Nq);
        else
          out[n] = sum_seq(&in[i + j * Nq], Nq2, &Mat[k * Nq], 1, Nq);
      }
}

/* element_operation_kernel_matvec_gradlike_dirXYZ_P<p>
 * (mod_dg_optr_kernel_opt1.F90.erb:166-219) as its comment states it: the
 * x, y and z derivatives of one field, out_x = Mat along x, out_y and out_z =
 * Mat_tr along y and z, left-to-right sums.  The reference's z loop
 * (:208-213) starts its sum with in(i,1,k)*Mat_tr(1,j) instead of
 * in(i,j,1)*Mat_tr(1,k); that slip is not reproduced here (the tests show
 * the difference on the reference's own output). */
void orc_gradlike(int Nq, const double *Mat, const double *Mat_tr, const double *in,
                  double *out_x, double *out_y, double *out_z) {
  orc_matvec_dir(Nq, 0, Mat, in, out_x);
  orc_matvec_dir(Nq, 1, Mat_tr, in, out_y);
  orc_matvec_dir(Nq, 2, Mat_tr, in, out_z);
}

/* element_operation_kernel_matvec_divlike_dirXYZ_var5_P<p>
 * (mod_dg_optr_kernel_opt1.F90.erb:347-465): divlike_dirXYZ in OPT1's order
 * on five consecutive fields v = 1..5 of one element (in/out (Np,5)).  The
 * reference's y loop writes out_y at jkv = j + (kv-1)*Nq*5 (:400), which
 * leaves most of out_y unwritten and runs past its end for kv > Nq; the
 * intended index j + (kv-1)*Nq is used here. */
void orc_divlike_var5(int Nq, const double *Mat, const double *Mat_tr, const double *in_x,
                      const double *in_y, const double *in_z, double *out_x, double *out_y,
                      double *out_z) {
  long Np = (long)Nq * Nq * Nq;
  for (int v = 0; v < 5; ++v)
    orc_divlike(Nq, ORC_OPT1, Mat, Mat_tr, in_x + v * Np, in_y + v * Np, in_z + v * Np,
                out_x + v * Np, out_y + v * Np, out_z + v * Np);
}

/* tensorprod_Lift_hexahedral for one element (mod_dg_optr_kernel.f90:262-288,
 * mod_dg_optr_kernel_opt1.F90.erb:134-157; both use the same order) */
void orc_lift(int Nq, const double *Lift, const double *in, double *out) {
  int Nq2 = Nq * Nq;
  long Np = (long)Nq2 * Nq;
  for (int k = 0; k < Nq; ++k)
    for (int j = 0; j < Nq; ++j)
      for (int i = 0; i < Nq; ++i) {
        long n = i + j * Nq + (long)k * Nq2;
        out[n] = Lift[n + 0 * Np] * in[i + k * Nq + 0 * Nq2] +
                 Lift[n + 1 * Np] * in[j + k * Nq + 1 * Nq2] +
                 Lift[n + 2 * Np] * in[i + k * Nq + 2 * Nq2] +
                 Lift[n + 3 * Np] * in[j + k * Nq + 3 * Nq2] +
                 Lift[n + 4 * Np] * in[i + j * Nq + 4 * Nq2] +
                 Lift[n + 5 * Np] * in[i + j * Nq + 5 * Nq2];
      }
}

/* cal_dqdt, mod_advect3d_eq.f90:146-201 */
void orc_cal_dqdt(double *dqdt, const double *q, const double *u,
                  const double *v, const double *w, const double *flux_bnd,
                  const double *D1D, const double *D1D_tr, const double *Lift,
                  const double *Escale, int Nq, int Ne, int typeid) {
  long Np = (long)Nq * Nq * Nq;
  int NfpTot = 6 * Nq * Nq;
  long NpNe = Np * Ne;
#pragma omp parallel
  {
    double *buf = (double *)malloc(sizeof(double) * 7 * Np);
    double *fx = buf, *fy = buf + Np, *fz = buf + 2 * Np;
    double *dx = buf + 3 * Np, *dy = buf + 4 * Np, *dz = buf + 5 * Np;
    double *lf = buf + 6 * Np;
#pragma omp for schedule(static)
    for (long ke = 0; ke < Ne; ++ke) {
      const double *qe = q + ke * Np;
      for (long n = 0; n < Np; ++n) {
        fx[n] = qe[n] * u[n + ke * Np];
        fy[n] = qe[n] * v[n + ke * Np];
        fz[n] = qe[n] * w[n + ke * Np];
      }
      orc_divlike(Nq, typeid, D1D, D1D_tr, fx, fy, fz, dx, dy, dz);
      orc_lift(Nq, Lift, flux_bnd + ke * NfpTot, lf);
      for (long n = 0; n < Np; ++n)
        dqdt[n + ke * Np] = -(Escale[n + ke * Np] * dx[n] +
                              Escale[n + ke * Np + NpNe] * dy[n] +
                              Escale[n + ke * Np + 2 * NpNe] * dz[n] + lf[n]);
    }
    free(buf);
  }
}

/* advect3d_eq_cal_tend, mod_advect3d_eq.f90:41-85.  ebnd_flux is a caller
 * workspace of NfpTot*Ne doubles (an automatic array in the reference). */
void orc_cal_tend(double *dqdt, const double *q, const double *u,
                  const double *v, const double *w, const double *D1D,
                  const double *D1D_tr, const double *Lift, const int *VMapM,
                  const int *VMapP, const double *normal_fn,
                  const double *Escale, const double *Fscale, int Nq, int Ne,
                  int typeid, double *ebnd_flux) {
  int Np = Nq * Nq * Nq, NfpTot = 6 * Nq * Nq;
  orc_elembnd_flux(ebnd_flux, q, u, v, w, VMapM, VMapP, normal_fn, Fscale, Np,
                   NfpTot, Ne);
  orc_cal_dqdt(dqdt, q, u, v, w, ebnd_flux, D1D, D1D_tr, Lift, Escale, Nq, Ne,
               typeid);
}

/* Time loop of main.f90:51-80: per step q0 <- q, then per stage update_halo,
 * cal_tend and q = rk_a*q0 + rk_b*(q + dt*dqdt).  q is (Np,NeA) with its halo
 * of u,v,w already filled by the caller.  If minmax is non-NULL, it receives
 * (min,max) of the owned q after every step (2*nstep doubles).  Returns the
 * wall time of the loop in seconds. */
double orc_run(double *q, const double *u, const double *v, const double *w,
               const double *D1D, const double *D1D_tr, const double *Lift,
               const int *VMapM, const int *VMapP, const double *normal_fn,
               const double *Escale, const double *Fscale,
               const int *halo_src_map, int NhaloNode, int Nq, int Ne, int NeA,
               int typeid, double dt, int nstep, double *minmax) {
  long Np = (long)Nq * Nq * Nq;
  int NfpTot = 6 * Nq * Nq;
  double *q0 = (double *)malloc(sizeof(double) * Np * NeA);
  double *dqdt = (double *)malloc(sizeof(double) * Np * NeA);
  double *ebnd = (double *)malloc(sizeof(double) * (long)NfpTot * Ne);
#ifdef _OPENMP
  double t0 = omp_get_wtime();
#endif
  for (int istep = 0; istep < nstep; ++istep) {
#pragma omp parallel for schedule(static)
    for (long ke = 0; ke < Ne; ++ke)
      memcpy(q0 + ke * Np, q + ke * Np, sizeof(double) * Np);
    for (int s = 0; s < 3; ++s) {
      orc_update_halo(q, (int)Np, Ne, NhaloNode, halo_src_map);
      orc_cal_tend(dqdt, q, u, v, w, D1D, D1D_tr, Lift, VMapM, VMapP,
                   normal_fn, Escale, Fscale, Nq, Ne, typeid, ebnd);
      const double a = rk_a[s], b = rk_b[s];
#pragma omp parallel for schedule(static)
      for (long ke = 0; ke < Ne; ++ke)
        for (long n = ke * Np; n < (ke + 1) * Np; ++n)
          q[n] = a * q0[n] + b * (q[n] + dt * dqdt[n]);
    }
    if (minmax) {
      double mn = q[0], mx = q[0];
      for (long n = 1; n < Np * Ne; ++n) {
        if (q[n] < mn) mn = q[n];
        if (q[n] > mx) mx = q[n];
      }
      minmax[2 * istep] = mn;
      minmax[2 * istep + 1] = mx;
    }
  }
  double el = 0.0;
#ifdef _OPENMP
  el = omp_get_wtime() - t0;
#endif
  free(q0);
  free(dqdt);
  free(ebnd);
  return el;
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
