/*
 * dgk.h -- C-ABI of the MI355X (gfx950) DG advection operator library.
 *
 * The library replaces the operator-kernel layer of SCALE-DG's 3D advection
 * extraction (Masahiro-Ren/scale-dg-kernel-extraction) with hand-written HIP
 * kernels.  A Fortran host binds these entry points with iso_c_binding and
 * selects them with DGOptrKernel_OptType = "HIP".
 *
 * Conventions (those of the reference, mod_advect3d_eq.f90:41-66):
 *   - arrays are Fortran column-major and passed as plain pointers;
 *   - fields are q(Np,NeA) with NeA = Ne + Nhalo; the halo slots
 *     q(Np*Ne+1 : Np*Ne+NhaloNode) are filled from halo_src_map;
 *   - VMapM/VMapP(NfpTot,Ne) and halo_src_map(NhaloNode) hold 1-based int32
 *     flat indices into q(Np*NeA);
 *   - real numbers are IEEE double (RP = real64, mod_common.f90:23).
 * Host pointers unless a function says "device".  Every function returns
 * DGK_OK (0) or a DGK_ERR_* code; dgk_last_error() describes the last failure
 * of the calling thread.  Calls on one solver must come from one host thread.
 */
#ifndef DGK_H
#define DGK_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DGK_OK 0
#define DGK_ERR_ARG 1        /* invalid argument or size                  */
#define DGK_ERR_OPTTYPE 2    /* unknown DGOptrKernel_OptType              */
#define DGK_ERR_ORDER 3      /* polynomial order outside 1..15            */
#define DGK_ERR_IO 4         /* operator data file missing or malformed   */
#define DGK_ERR_HIP 5        /* HIP runtime error                         */
#define DGK_ERR_NOMEM 6      /* device allocation failed                  */
#define DGK_ERR_STATE 7      /* call not valid in the solver's state      */
#define DGK_ERR_NONFINITE 8  /* the state holds a NaN or an infinity      */

/* Operator evaluation types.  GENERAL and OPT1 are the reference's
 * (mod_dg_optr_kernel.f90:39-42); HIP is the device path added here.  The
 * summation order of a device run is GENERAL's or OPT1's. */
#define DGK_TYPE_GENERAL 1
#define DGK_TYPE_OPT1 2
#define DGK_TYPE_HIP 3

/* Solver flags (dgk_create / dgk_create_structured). */
#define DGK_FLAG_FMA 0x1           /* fused multiply-add contractions (not
                                      bit-identical to the reference)       */
#define DGK_FLAG_FORCE_TABLES 0x2  /* always use the per-node VMapM/VMapP,
                                      normal, Fscale and Escale tables       */
#define DGK_FLAG_NODE_KERNEL 0x4   /* thread-per-node stage kernel even where
                                      the pencil kernels apply (compressed
                                      connectivity) */
#define DGK_FLAG_FACTORED_LIFT 0x8 /* lift with the 1D vectors
                                      Lift(:,1,1,4) and Lift(:,1,1,2) (exact
                                      to ~1e-14 relative, not bit-identical) */
#define DGK_FLAG_CONST_VELOCITY 0x10 /* u,v,w uniform (checked by
                                      dgk_set_fields): only q and q0 stream,
                                      21.33 B per DOF-update instead of
                                      45.33 on average over a step; same
                                      results (pencil kernel only)          */

#define DGK_FLAG_MIXED 0x20         /* directional contractions in FP32:
                                      each line's cubic through its nodes
                                      0, Nq/3, 2Nq/3, Nq-1 is differentiated
                                      in FP64 and only the remainder is
                                      rounded to FP32 and contracted with
                                      the operator split into hi + lo FP32
                                      parts (two FP32 sums); a line whose
                                      remainder exceeds 1/16 of its largest
                                      value is contracted in FP64.  Lift,
                                      face-flux differences, residual and
                                      RK in FP64.  Implies the factored
                                      lift; pencil kernels only (compressed
                                      connectivity), no FMA flag */

#define DGK_FLAG_MFMA 0x40          /* directional contractions on the matrix
                                      cores: FP64 v_mfma_f64_16x16x4_f64,
                                      or with DGK_FLAG_MIXED FP32
                                      v_mfma_f32_16x16x4_f32 on each line's
                                      remainder after its degree-7
                                      interpolant through 8 of its nodes
                                      (interpolant and its derivative as
                                      FP64 MFMA products); face fluxes,
                                      lift, residual and RK in FP64.
                                      PolyOrder 15 with
                                      DGK_FLAG_CONST_VELOCITY on a compressed
                                      mesh; implies the factored lift; fused
                                      (not bit-identical) */

/* Fields addressable through dgk_get_field / dgk_field_ptr. */
#define DGK_FIELD_Q 0
#define DGK_FIELD_Q0 1
#define DGK_FIELD_DQDT 2
#define DGK_FIELD_U 3
#define DGK_FIELD_V 4
#define DGK_FIELD_W 5

typedef struct dgk_solver dgk_solver;

/* ---- status ------------------------------------------------------------ */
const char *dgk_last_error(void);
int dgk_version(void);

/* ---- selector ---------------------------------------------------------- */
/* Maps "GENERAL" -> 1, "OPT1" -> 2, "HIP" -> 3 (trailing blanks ignored).
 * Replaces dg_optr_kernel_setup (mod_dg_optr_kernel.f90:45-58), which only
 * prints on an unknown name; here an unknown name returns DGK_ERR_OPTTYPE. */
int dgk_optr_kernel_setup(const char *evaltype, int *type_id_out);

/* ---- element and mesh producers (host) -------------------------------- */
/* Reads operator_data/p<p>.dat: x1D(Nq), D1D(Nq,Nq), Lift_mat(Nq,Nq,Nq,6).
 * Replaces element_init's read (mod_mesh.f90:253-263). */
int dgk_read_operator(const char *dir, int p, double *x1D, double *D1D,
                      double *Lift_mat);

/* sizes[7] = {Nq, Np, Nfp, NfpTot, Ne, NeA, NhaloNode}
 * (mod_mesh.f90:96-100, 238-241). */
int dgk_mesh_sizes(int NeX, int NeY, int NeZ, int p, int32_t *sizes);

/* Builds the periodic structured hexahedral mesh tables in the reference's
 * numbering (mod_mesh.f90:74-193, 278-397): VMapM, VMapP(NfpTot,Ne),
 * normal_fn(NfpTot,Ne,3), Escale(Np,Ne,3), Fscale(NfpTot,Ne),
 * pos_en(Np,Ne,3), halo_src_map(NhaloNode).  Any output may be NULL. */
int dgk_mesh_setup(int NeX, int NeY, int NeZ, int p, double domx, double domy,
                   double domz, const double *x1D, int32_t *VMapM,
                   int32_t *VMapP, double *normal_fn, double *Escale,
                   double *Fscale, double *pos_en, int32_t *halo_src_map);

/* The same tables built on the device, straight into device arrays (any may
 * be NULL; x1D is a host array).  Replaces mesh_setup's table loops
 * (mod_mesh.f90:74-193) including the serial halo numbering (:171-190),
 * which each face node derives in closed form.  Synchronous on stream. */
int dgk_mesh_setup_device(int NeX, int NeY, int NeZ, int p, double domx,
                          double domy, double domz, const double *x1D,
                          int32_t *VMapM, int32_t *VMapP, double *normal_fn,
                          double *Escale, double *Fscale, double *pos_en,
                          int32_t *halo_src_map, void *stream);

/* ---- device-resident solver ------------------------------------------- */
/* Creates a solver from the reference's mesh-level arrays (the argument list
 * of advect3d_eq_cal_tend, mod_advect3d_eq.f90:41-66, plus the halo map of
 * update_halo, mod_mesh.f90:209-221).  The tables are verified on the host;
 * when they describe a periodic structured mesh with axis-aligned uniform
 * geometry they are compressed to a 6-neighbour table and per-face/per-axis
 * constants, otherwise the per-node tables are kept on the device.
 * type_id: DGK_TYPE_GENERAL or DGK_TYPE_OPT1 (summation order); DGK_TYPE_HIP
 * means OPT1's order. */
int dgk_create(dgk_solver **out, int Nq, int Ne, int NeA, int NfpTot,
               const double *D1D, const double *Lift_mat,
               const int32_t *VMapM, const int32_t *VMapP,
               const double *normal_fn, const double *Escale,
               const double *Fscale, const int32_t *halo_src_map,
               int NhaloNode, int type_id, int flags);

/* Same solver for the periodic NeX x NeY x NeZ mesh on [0,domx]x[0,domy]x
 * [0,domz] built directly in compressed form (no per-node tables). */
int dgk_create_structured(dgk_solver **out, int NeX, int NeY, int NeZ, int p,
                          double domx, double domy, double domz,
                          const double *D1D, const double *Lift_mat,
                          int type_id, int flags);

int dgk_destroy(dgk_solver *s);

/* Uploads q,u,v,w (Np,NeA) (owned parts used; halos are refilled on the
 * device, as main.f90:159-162 does after the initial condition).
 * Synchronous.  Fields above 16 MiB (here and in dgk_get_field, the
 * ensemble's transfers and the stateless tendency) go through two pinned
 * 64 MiB staging buffers per device, allocated on first use and held until
 * the process ends, filled and emptied by up to 8 host threads of the
 * library's own while the DMA of the other buffer runs. */
int dgk_set_fields(dgk_solver *s, const double *q, const double *u,
                   const double *v, const double *w);
/* Downloads one field (Np,NeA).  For q the halo slots hold the values of
 * the current owned q. */
int dgk_get_field(dgk_solver *s, int which, double *out);
/* Device address of a field (Np*NeA doubles) for zero-copy callers; for q
 * and q0 valid until the next stage or step. */
int dgk_field_ptr(dgk_solver *s, int which, double **dptr);

/* dqdt = advect3d_eq_cal_tend(q,u,v,w,...) on the device
 * (mod_advect3d_eq.f90:41-85), halos of q refreshed first. */
int dgk_cal_tend(dgk_solver *s);
/* One SSP-RK3 stage (stage = 1..3): update_halo(q), tendency and
 * q = rk_a(stage) q0 + rk_b(stage) (q + dt dqdt) fused in one launch; stage
 * 1 also makes q0 <- q (main.f90:53-73) by keeping the buffer it read as q0,
 * so the device addresses of q and q0 (dgk_field_ptr) change with every
 * stage. */
int dgk_rk_stage(dgk_solver *s, int stage, double dt);
/* One full time step (three stages) replayed from a captured hipGraph. */
int dgk_step(dgk_solver *s, double dt);
/* nstep time steps, each as three stage launches on the solver's stream (on
 * small meshes faster than dgk_step's graph replays; the same bits); if
 * minmax is non-NULL it receives (min,max) of the owned
 * q after every step, 2*nstep doubles (the main.f90:76-79 diagnostics), and
 * a state with a NaN or an infinity makes the call return DGK_ERR_NONFINITE
 * (naming the first such step) after all nstep steps and minmax are done; a
 * NaN shows as min and max of that step.  Without minmax nothing is checked. */
int dgk_run(dgk_solver *s, double dt, int nstep, double *minmax);
int dgk_synchronize(dgk_solver *s);
/* what: 0 = connectivity mode (1 compressed, 0 tables), 1 = Ne, 2 = Np,
 * 3 = NeA, 4 = grid size of the stage kernel, 5 = block size, 6 = stage
 * kernel kind (0 thread-per-node, 1 pencil), 7 = LDS bytes per workgroup,
 * 8 = tuning variant of the stage kernel, 9 = tiled element order (1) or
 * linear (0). */
int dgk_query(dgk_solver *s, int what, int64_t *value);
/* Tuning variants of the stage kernel (code shapes with identical results):
 * *n receives their number; dgk_set_variant selects one (-1 = the built-in
 * default for this Nq).  Used by the autotuner. */
int dgk_num_variants(dgk_solver *s, int *n);
int dgk_set_variant(dgk_solver *s, int variant);
/* Times `nstage` consecutive stage launches with HIP events on the solver's
 * stream; returns the mean milliseconds per stage launch.  The launches are
 * whole SSP-RK3 steps (stage 1, 2, 3, 1, ...): nstage must be a multiple of
 * 3, and the state advances by nstage/3 steps. */
int dgk_time_stages(dgk_solver *s, double dt, int nstage, double *ms_per_stage);
/* Autotuner: times every tuning variant (workgroup shape, element order,
 * store flavour; results are the same bits) over nrep stages (rounded up to
 * whole steps) at this dt, keeps the fastest and leaves the state as it was.  variant_out and
 * ms_out (average stage time of the winner) may be NULL. */
int dgk_autotune(dgk_solver *s, double dt, int nrep, int *variant_out,
                 double *ms_out);

/* ---- diagnostics (device reductions) ------------------------------------ */
/* out[8] for the owned state q, in a fixed reduction order:
 *   0 min q, 1 max q (the main.f90:76-79 line),
 *   2 mass = sum_e J_e sum_ijk w_i w_j w_k q, 3 the same sum of |q|
 *     (w: the LGL weights w1D[Nq], J = 1/(Es1 Es2 Es3) = element volume/8),
 *   4 sum q^2 over owned DOF,
 *   5 sum (q-ref)^2, 6 sum ref^2, 7 max |q-ref| for a device field
 *     ref(Np,Ne) (e.g. dgk_field_ptr of another solver), or 0 if ref is NULL.
 * Relative L2 error = sqrt(out[5]/out[6]). */
int dgk_diagnostics(dgk_solver *s, const double *w1D, const double *ref,
                    double *out);

/* dgk_diagnostics that allocates nothing.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The same
 * reductions in the same order (the same bits) as dgk_diagnostics, but the
 * eight results go to out_dev, a device array of 8 doubles, in stream order
 * on the solver's stream, and the call does not wait for them.  The
 * per-block partial sums live in a file-scope device array of 1024 x 8
 * doubles (64 KiB per GPU, held until the process ends; the reduction never
 * uses more than 1024 blocks, whatever the field's size).  Every call on one
 * GPU shares that array: calls that overlap on one GPU race. */
int dgk_diagnostics_noalloc(dgk_solver *s, const double *w1D, const double *ref,
                            double *out_dev);

/* ---- stateless drop-in of the mesh-level tendency ----------------------- */
/* Same arguments and semantics as advect3d_eq_cal_tend
 * (mod_advect3d_eq.f90:41-45); host arrays; q,u,v,w must carry their halos.
 * The tendency is evaluated on the device in OPT1's summation order unless
 * type_id is DGK_TYPE_GENERAL. */
int dgk_advect3d_eq_cal_tend(double *dqdt, const double *q, const double *u,
                             const double *v, const double *w,
                             const double *D1D, const double *D1D_tr,
                             const double *Lift_mat, const int32_t *VMapM,
                             const int32_t *VMapP, const double *normal_fn,
                             const double *Escale, const double *Fscale,
                             int Nq, int Np, int NfpTot, int Ne, int NeA,
                             int type_id);

/* ---- batched element operators (device arrays) ------------------------- */
/* tensorprod_divlike_dirXYZ (mod_dg_optr_kernel.f90:62-136) over nelem
 * element blocks: in_x/in_y/in_z/out_* are (Np,nelem), Mat/Mat_tr (Nq,Nq).
 * type_id selects GENERAL or OPT1 summation.  stream: hipStream_t or NULL. */
int dgk_divlike_dirXYZ_batched(int Nq, int nelem, int type_id,
                               const double *Mat, const double *Mat_tr,
                               const double *in_x, const double *in_y,
                               const double *in_z, double *out_x,
                               double *out_y, double *out_z, void *stream);
/* tensorprod_Lift_hexahedral (mod_dg_optr_kernel.f90:139-190) over nelem
 * blocks: in (NfpTot,nelem), out (Np,nelem), Lift (Nq,Nq,Nq,6). */
int dgk_lift_hexahedral_batched(int Nq, int nelem, const double *Lift,
                                const double *in, double *out, void *stream);

/* The rest of the OPT1 operator family (mod_dg_optr_kernel_opt1.F90.erb),
 * over nelem element blocks (device arrays, (Np,nelem) each):
 * element_operation_kernel_matvec_dir{X,Y,Z}_P<p> (:51-129): dir 0
 * out(i,jk) = Mat(i,l) in(l,jk), dir 1 out(i,j,k) = in(i,l,k) Mat(l,j), dir 2
 * out(i,j,k) = in(i,j,l) Mat(l,k) (Mat = Mat_x, or the transposed Mat_y,
 * Mat_z as the reference's callers pass them); left-to-right sums. */
int dgk_matvec_dir_batched(int Nq, int nelem, int dir, const double *Mat,
                           const double *in, double *out, void *stream);
/* element_operation_kernel_matvec_gradlike_dirXYZ_P<p> (:166-219) in its
 * stated form: the x, y, z derivatives of one field (in(Np,nelem)).  The
 * reference's z loop begins with in(i,1,k)*Mat_tr(1,j) (:208); here it is
 * in(i,j,1)*Mat_tr(1,k), so out_z differs from the reference's. */
int dgk_gradlike_dirXYZ_batched(int Nq, int nelem, const double *Mat,
                                const double *Mat_tr, const double *in,
                                double *out_x, double *out_y, double *out_z,
                                void *stream);
/* element_operation_kernel_matvec_divlike_dirXYZ_var5_P<p> (:347-465):
 * divlike_dirXYZ in OPT1's order on 5 fields per element (in/out
 * (Np,5,nelem)).  The reference's y loop indexes its output with
 * j + (kv-1)*Nq*5 (:400); here with the intended j + (kv-1)*Nq. */
int dgk_divlike_dirXYZ_var5_batched(int Nq, int nelem, const double *Mat,
                                    const double *Mat_tr, const double *in_x,
                                    const double *in_y, const double *in_z,
                                    double *out_x, double *out_y,
                                    double *out_z, void *stream);

/* ---- batched ensemble ---------------------------------------------------- */
/* n independent runs of the reference program (main.f90:51-80), each on its
 * own periodic NeX x NeY x NeZ mesh of order p with its own velocity field,
 * initial state and dt, advanced together: one persistent launch per RK
 * stage over all instances, the three stages of a step captured in one
 * hipGraph.  Every instance gives the same bits as a dgk_create_structured
 * solver with the same flags (OPT1 summation order).  PolyOrder 1..15.
 * dom: n triples of domain lengths, or NULL for unit cubes.
 * operator_dir: the directory of the reference's pN.dat files.  flags:
 * DGK_FLAG_FMA, DGK_FLAG_FACTORED_LIFT, DGK_FLAG_CONST_VELOCITY,
 * DGK_FLAG_MIXED (implies the factored lift; not with DGK_FLAG_FMA). */
typedef struct dgk_ensemble dgk_ensemble;
int dgk_ensemble_create(dgk_ensemble **out, int n, const int32_t *p,
                        const int32_t *NeX, const int32_t *NeY,
                        const int32_t *NeZ, const double *dom,
                        const char *operator_dir, int type_id, int flags);
int dgk_ensemble_destroy(dgk_ensemble *s);
/* instance i: q,u,v,w (Np,Ne) host arrays (main.f90:135-157) */
int dgk_ensemble_set_fields(dgk_ensemble *s, int i, const double *q,
                            const double *u, const double *v, const double *w);
int dgk_ensemble_get_field(dgk_ensemble *s, int i, double *q);
/* dt[n]: time step of every instance (the namelist's dt, main.f90:103-107) */
int dgk_ensemble_set_dt(dgk_ensemble *s, const double *dt);
/* one SSP-RK3 step of every instance (graph replay, asynchronous) */
int dgk_ensemble_step(dgk_ensemble *s);
/* nstep steps, then waits for the device */
int dgk_ensemble_run(dgk_ensemble *s, int nstep);
/* average device time of one stage launch over nstage launches: whole
 * SSP-RK3 steps (nstage a multiple of 3), which advance the state */
int dgk_ensemble_time_stages(dgk_ensemble *s, int nstage, double *ms_per_stage);
/* Chooses the tile shape and the grid split from measured times (nrep
 * stage-1 launches per trial; the state is left unchanged): for each
 * workgroup width the largest order allows (192 up to PolyOrder 12, 256),
 * every order class tries its candidate elements per item (as many as fill
 * the workgroup, about half, one) timed alone; the grid is split between the
 * classes by those times, by DOF and by the geometric mean of the two, the
 * whole stage is timed for each, the fastest width and split are kept, and
 * single units of eight workgroups move between the largest class and the
 * others while the whole stage gets faster.  Results are bit-identical for
 * every shape and split.
 * ms_per_class[number of classes]: each class's time alone in the chosen
 * shape, or NULL. */
int dgk_ensemble_tune(dgk_ensemble *s, int nrep, double *ms_per_class);
/* Sets the tile shape by hand: threads per workgroup (192 or 256; 192 only
 * without an instance of PolyOrder 13..15; 0: the default for the largest
 * order) and elems_per_item[number of classes] (0 or NULL: as many as fill
 * the workgroup).  Drops a previous tuning (the grid is split by DOF). */
int dgk_ensemble_set_shape(dgk_ensemble *s, int threads, const int32_t *elems_per_item);
/* what: 0 total DOF, 1 work items per stage, 2 workgroups, 3 LDS bytes,
 * 4 classes of equal order, 5 tuned (0/1), 6 threads per workgroup (192, or
 * 256 with an instance of PolyOrder 13..15, unless tuned or set), 16+c Nq of
 * class c, 32+c workgroups of class c, 48+c elements per item of class c */
int dgk_ensemble_query(dgk_ensemble *s, int what, int64_t *value);
/* dgk_diagnostics for instance i: out[8] over its owned state, against the
 * device field ref (Np,Ne) or NULL; w1D its LGL weights (Nq). */
int dgk_ensemble_diagnostics(dgk_ensemble *s, int i, const double *w1D, const double *ref,
                             double *out);

#ifdef __cplusplus
}
#endif
#endif /* DGK_H */
