!-------------------------------------------------------------------------------
!> mod_dgk_hip: iso_c_binding interfaces to libdgk.so (include/dgk.h), the
!! MI355X operator library behind DGOptrKernel_OptType = "HIP".
!!
!! Arrays are passed by reference in the reference's Fortran layouts
!! (q(Np,NeA), VMapM(NfpTot,Ne), normal_fn(NfpTot,Ne,3), ...), integers are
!! default-kind (int32) and reals real64, exactly as advect3d_eq_cal_tend
!! receives them (mod_advect3d_eq.f90:41-66).  Every function returns 0 on
!! success; dgk_error_message() gives the text of the last failure.
!-------------------------------------------------------------------------------
module mod_dgk_hip
  use, intrinsic :: iso_c_binding
  implicit none
  private

  integer, public, parameter :: DGK_OK = 0
  integer, public, parameter :: DGK_ERR_ARG = 1, DGK_ERR_OPTTYPE = 2, DGK_ERR_ORDER = 3, &
       DGK_ERR_IO = 4, DGK_ERR_HIP = 5, DGK_ERR_NOMEM = 6, DGK_ERR_STATE = 7, &
       DGK_ERR_NONFINITE = 8
  integer, public, parameter :: DGK_TYPE_GENERAL = 1, DGK_TYPE_OPT1 = 2, DGK_TYPE_HIP = 3
  integer, public, parameter :: DGK_FLAG_FMA = 1, DGK_FLAG_FORCE_TABLES = 2, &
       DGK_FLAG_NODE_KERNEL = 4, DGK_FLAG_FACTORED_LIFT = 8, DGK_FLAG_CONST_VELOCITY = 16, &
       DGK_FLAG_MIXED = 32, DGK_FLAG_MFMA = 64
  integer, public, parameter :: DGK_FIELD_Q = 0, DGK_FIELD_Q0 = 1, DGK_FIELD_DQDT = 2, &
       DGK_FIELD_U = 3, DGK_FIELD_V = 4, DGK_FIELD_W = 5

  public :: dgk_version, dgk_optr_kernel_setup, dgk_read_operator, dgk_mesh_sizes
  public :: dgk_mesh_setup, dgk_mesh_setup_device
  public :: dgk_create, dgk_create_structured, dgk_destroy
  public :: dgk_set_fields, dgk_get_field, dgk_field_ptr, dgk_cal_tend, dgk_rk_stage, dgk_step
  public :: dgk_run, dgk_synchronize, dgk_query, dgk_num_variants, dgk_set_variant
  public :: dgk_time_stages, dgk_advect3d_eq_cal_tend, dgk_error_message
  public :: dgk_autotune, dgk_diagnostics, dgk_diagnostics_noalloc
  public :: dgk_divlike_dirXYZ_batched, dgk_lift_hexahedral_batched, dgk_matvec_dir_batched
  public :: dgk_gradlike_dirXYZ_batched, dgk_divlike_dirXYZ_var5_batched
  public :: dgk_ensemble_create, dgk_ensemble_destroy, dgk_ensemble_set_fields
  public :: dgk_ensemble_get_field, dgk_ensemble_set_dt, dgk_ensemble_step, dgk_ensemble_run
  public :: dgk_ensemble_time_stages, dgk_ensemble_tune, dgk_ensemble_query
  public :: dgk_ensemble_set_shape
  public :: dgk_ensemble_diagnostics

  interface
    integer(c_int) function dgk_optr_kernel_setup(evaltype, type_id) bind(C, name="dgk_optr_kernel_setup")
      import :: c_int, c_char
      character(kind=c_char), intent(in) :: evaltype(*)
      integer(c_int), intent(out) :: type_id
    end function

    integer(c_int) function dgk_read_operator(dir, p, x1D, D1D, Lift_mat) bind(C, name="dgk_read_operator")
      import :: c_int, c_char, c_double
      character(kind=c_char), intent(in) :: dir(*)
      integer(c_int), value :: p
      real(c_double), intent(out) :: x1D(*), D1D(*), Lift_mat(*)
    end function

    integer(c_int) function dgk_mesh_sizes(NeX, NeY, NeZ, p, sizes) bind(C, name="dgk_mesh_sizes")
      import :: c_int
      integer(c_int), value :: NeX, NeY, NeZ, p
      integer(c_int), intent(out) :: sizes(7)
    end function

    integer(c_int) function dgk_mesh_setup(NeX, NeY, NeZ, p, domx, domy, domz, x1D, &
         VMapM, VMapP, normal_fn, Escale, Fscale, pos_en, halo_src_map) bind(C, name="dgk_mesh_setup")
      import :: c_int, c_double
      integer(c_int), value :: NeX, NeY, NeZ, p
      real(c_double), value :: domx, domy, domz
      real(c_double), intent(in) :: x1D(*)
      integer(c_int), intent(out) :: VMapM(*), VMapP(*), halo_src_map(*)
      real(c_double), intent(out) :: normal_fn(*), Escale(*), Fscale(*), pos_en(*)
    end function

    integer(c_int) function dgk_create(solver, Nq, Ne, NeA, NfpTot, D1D, Lift_mat, VMapM, VMapP, &
         normal_fn, Escale, Fscale, halo_src_map, NhaloNode, type_id, flags) bind(C, name="dgk_create")
      import :: c_int, c_double, c_ptr
      type(c_ptr), intent(out) :: solver
      integer(c_int), value :: Nq, Ne, NeA, NfpTot, NhaloNode, type_id, flags
      real(c_double), intent(in) :: D1D(*), Lift_mat(*), normal_fn(*), Escale(*), Fscale(*)
      integer(c_int), intent(in) :: VMapM(*), VMapP(*), halo_src_map(*)
    end function

    integer(c_int) function dgk_create_structured(solver, NeX, NeY, NeZ, p, domx, domy, domz, &
         D1D, Lift_mat, type_id, flags) bind(C, name="dgk_create_structured")
      import :: c_int, c_double, c_ptr
      type(c_ptr), intent(out) :: solver
      integer(c_int), value :: NeX, NeY, NeZ, p, type_id, flags
      real(c_double), value :: domx, domy, domz
      real(c_double), intent(in) :: D1D(*), Lift_mat(*)
    end function

    integer(c_int) function dgk_destroy(solver) bind(C, name="dgk_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: solver
    end function

    integer(c_int) function dgk_set_fields(solver, q, u, v, w) bind(C, name="dgk_set_fields")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: solver
      real(c_double), intent(in) :: q(*), u(*), v(*), w(*)
    end function

    integer(c_int) function dgk_get_field(solver, which, f) bind(C, name="dgk_get_field")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: solver
      integer(c_int), value :: which
      real(c_double), intent(out) :: f(*)
    end function

    integer(c_int) function dgk_cal_tend(solver) bind(C, name="dgk_cal_tend")
      import :: c_int, c_ptr
      type(c_ptr), value :: solver
    end function

    integer(c_int) function dgk_rk_stage(solver, stage, dt) bind(C, name="dgk_rk_stage")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: solver
      integer(c_int), value :: stage
      real(c_double), value :: dt
    end function

    integer(c_int) function dgk_step(solver, dt) bind(C, name="dgk_step")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: solver
      real(c_double), value :: dt
    end function

    integer(c_int) function dgk_run(solver, dt, nstep, minmax) bind(C, name="dgk_run")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: solver
      real(c_double), value :: dt
      integer(c_int), value :: nstep
      real(c_double), intent(out) :: minmax(*)
    end function

    integer(c_int) function dgk_synchronize(solver) bind(C, name="dgk_synchronize")
      import :: c_int, c_ptr
      type(c_ptr), value :: solver
    end function

    integer(c_int) function dgk_advect3d_eq_cal_tend(dqdt, q, u, v, w, D1D, D1D_tr, Lift_mat, &
         VMapM, VMapP, normal_fn, Escale, Fscale, Nq, Np, NfpTot, Ne, NeA, type_id) &
         bind(C, name="dgk_advect3d_eq_cal_tend")
      import :: c_int, c_double
      real(c_double), intent(out) :: dqdt(*)
      real(c_double), intent(in) :: q(*), u(*), v(*), w(*), D1D(*), D1D_tr(*), Lift_mat(*)
      integer(c_int), intent(in) :: VMapM(*), VMapP(*)
      real(c_double), intent(in) :: normal_fn(*), Escale(*), Fscale(*)
      integer(c_int), value :: Nq, Np, NfpTot, Ne, NeA, type_id
    end function

    integer(c_int) function dgk_autotune(solver, dt, nrep, variant, ms) bind(C, name="dgk_autotune")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: solver
      real(c_double), value :: dt
      integer(c_int), value :: nrep
      integer(c_int), intent(out) :: variant
      real(c_double), intent(out) :: ms
    end function

    ! out(8): min, max, mass, |q| mass, sum q^2, sum (q-ref)^2, sum ref^2, max |q-ref|;
    ! ref is a device address (c_null_ptr for none)
    integer(c_int) function dgk_diagnostics(solver, w1D, ref, out) bind(C, name="dgk_diagnostics")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: solver
      real(c_double), intent(in) :: w1D(*)
      type(c_ptr), value :: ref
      real(c_double), intent(out) :: out(8)
    end function

    ! the ensemble: n independent runs advanced together
    integer(c_int) function dgk_ensemble_create(ens, n, p, NeX, NeY, NeZ, dom, operator_dir, &
         type_id, flags) bind(C, name="dgk_ensemble_create")
      import :: c_int, c_ptr, c_char
      type(c_ptr), intent(out) :: ens
      integer(c_int), value :: n, type_id, flags
      integer(c_int), intent(in) :: p(*), NeX(*), NeY(*), NeZ(*)
      type(c_ptr), value :: dom
      character(kind=c_char), intent(in) :: operator_dir(*)
    end function

    integer(c_int) function dgk_ensemble_destroy(ens) bind(C, name="dgk_ensemble_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ens
    end function

    integer(c_int) function dgk_ensemble_set_fields(ens, i, q, u, v, w) &
         bind(C, name="dgk_ensemble_set_fields")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ens
      integer(c_int), value :: i
      real(c_double), intent(in) :: q(*), u(*), v(*), w(*)
    end function

    integer(c_int) function dgk_ensemble_get_field(ens, i, q) bind(C, name="dgk_ensemble_get_field")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ens
      integer(c_int), value :: i
      real(c_double), intent(out) :: q(*)
    end function

    integer(c_int) function dgk_ensemble_set_dt(ens, dt) bind(C, name="dgk_ensemble_set_dt")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ens
      real(c_double), intent(in) :: dt(*)
    end function

    integer(c_int) function dgk_ensemble_run(ens, nstep) bind(C, name="dgk_ensemble_run")
      import :: c_int, c_ptr
      type(c_ptr), value :: ens
      integer(c_int), value :: nstep
    end function

    integer(c_int) function dgk_ensemble_tune(ens, nrep, ms) bind(C, name="dgk_ensemble_tune")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ens
      integer(c_int), value :: nrep
      real(c_double), intent(out) :: ms(*)
    end function

    ! tile shape by hand: threads 192 or 256 (0: default), elements per item
    ! of each order class (0: as many as fill the workgroup)
    integer(c_int) function dgk_ensemble_set_shape(ens, threads, epb) &
         bind(C, name="dgk_ensemble_set_shape")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ens
      integer(c_int), value :: threads
      integer(c_int32_t), intent(in) :: epb(*)
    end function

    ! dgk_diagnostics for instance i (0-based) of the ensemble
    integer(c_int) function dgk_ensemble_diagnostics(ens, i, w1D, ref, out) &
         bind(C, name="dgk_ensemble_diagnostics")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ens
      integer(c_int), value :: i
      real(c_double), intent(in) :: w1D(*)
      type(c_ptr), value :: ref
      real(c_double), intent(out) :: out(8)
    end function

    type(c_ptr) function dgk_last_error() bind(C, name="dgk_last_error")
      import :: c_ptr
    end function

    integer(c_int) function dgk_version() bind(C, name="dgk_version")
      import :: c_int
    end function

    ! the mesh tables built on the device (device arrays; any may be c_null_ptr)
    integer(c_int) function dgk_mesh_setup_device(NeX, NeY, NeZ, p, domx, domy, domz, x1D, &
         VMapM, VMapP, normal_fn, Escale, Fscale, pos_en, halo_src_map, stream) &
         bind(C, name="dgk_mesh_setup_device")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: NeX, NeY, NeZ, p
      real(c_double), value :: domx, domy, domz
      real(c_double), intent(in) :: x1D(*)
      type(c_ptr), value :: VMapM, VMapP, normal_fn, Escale, Fscale, pos_en, halo_src_map, stream
    end function

    ! device address of a field (Np*NeA doubles); q and q0 move with every stage
    integer(c_int) function dgk_field_ptr(solver, which, dptr) bind(C, name="dgk_field_ptr")
      import :: c_int, c_ptr
      type(c_ptr), value :: solver
      integer(c_int), value :: which
      type(c_ptr), intent(out) :: dptr
    end function

    integer(c_int) function dgk_query(solver, what, val) bind(C, name="dgk_query")
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: solver
      integer(c_int), value :: what
      integer(c_int64_t), intent(out) :: val
    end function

    integer(c_int) function dgk_num_variants(solver, n) bind(C, name="dgk_num_variants")
      import :: c_int, c_ptr
      type(c_ptr), value :: solver
      integer(c_int), intent(out) :: n
    end function

    integer(c_int) function dgk_set_variant(solver, variant) bind(C, name="dgk_set_variant")
      import :: c_int, c_ptr
      type(c_ptr), value :: solver
      integer(c_int), value :: variant
    end function

    ! nstage a multiple of 3 (whole steps; the state advances)
    integer(c_int) function dgk_time_stages(solver, dt, nstage, ms) bind(C, name="dgk_time_stages")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: solver
      real(c_double), value :: dt
      integer(c_int), value :: nstage
      real(c_double), intent(out) :: ms
    end function

    ! dgk_diagnostics into the device array out_dev(8), without allocating or waiting
    integer(c_int) function dgk_diagnostics_noalloc(solver, w1D, ref, out_dev) &
         bind(C, name="dgk_diagnostics_noalloc")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: solver
      real(c_double), intent(in) :: w1D(*)
      type(c_ptr), value :: ref, out_dev
    end function

    ! the reference's operator kernels over nelem element blocks (device arrays)
    integer(c_int) function dgk_divlike_dirXYZ_batched(Nq, nelem, type_id, Mat, Mat_tr, &
         in_x, in_y, in_z, out_x, out_y, out_z, stream) bind(C, name="dgk_divlike_dirXYZ_batched")
      import :: c_int, c_ptr
      integer(c_int), value :: Nq, nelem, type_id
      type(c_ptr), value :: Mat, Mat_tr, in_x, in_y, in_z, out_x, out_y, out_z, stream
    end function

    integer(c_int) function dgk_lift_hexahedral_batched(Nq, nelem, Lift, fin, fout, stream) &
         bind(C, name="dgk_lift_hexahedral_batched")
      import :: c_int, c_ptr
      integer(c_int), value :: Nq, nelem
      type(c_ptr), value :: Lift, fin, fout, stream
    end function

    integer(c_int) function dgk_matvec_dir_batched(Nq, nelem, dir, Mat, fin, fout, stream) &
         bind(C, name="dgk_matvec_dir_batched")
      import :: c_int, c_ptr
      integer(c_int), value :: Nq, nelem, dir
      type(c_ptr), value :: Mat, fin, fout, stream
    end function

    integer(c_int) function dgk_gradlike_dirXYZ_batched(Nq, nelem, Mat, Mat_tr, fin, &
         out_x, out_y, out_z, stream) bind(C, name="dgk_gradlike_dirXYZ_batched")
      import :: c_int, c_ptr
      integer(c_int), value :: Nq, nelem
      type(c_ptr), value :: Mat, Mat_tr, fin, out_x, out_y, out_z, stream
    end function

    integer(c_int) function dgk_divlike_dirXYZ_var5_batched(Nq, nelem, Mat, Mat_tr, &
         in_x, in_y, in_z, out_x, out_y, out_z, stream) &
         bind(C, name="dgk_divlike_dirXYZ_var5_batched")
      import :: c_int, c_ptr
      integer(c_int), value :: Nq, nelem
      type(c_ptr), value :: Mat, Mat_tr, in_x, in_y, in_z, out_x, out_y, out_z, stream
    end function

    integer(c_int) function dgk_ensemble_step(ens) bind(C, name="dgk_ensemble_step")
      import :: c_int, c_ptr
      type(c_ptr), value :: ens
    end function

    integer(c_int) function dgk_ensemble_time_stages(ens, nstage, ms) &
         bind(C, name="dgk_ensemble_time_stages")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ens
      integer(c_int), value :: nstage
      real(c_double), intent(out) :: ms
    end function

    integer(c_int) function dgk_ensemble_query(ens, what, val) bind(C, name="dgk_ensemble_query")
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: ens
      integer(c_int), value :: what
      integer(c_int64_t), intent(out) :: val
    end function
  end interface

contains

  !> Text of the last library error of this thread.
  function dgk_error_message() result(msg)
    character(len=512) :: msg
    character(kind=c_char), pointer :: p(:)
    integer :: i
    msg = ""
    call c_f_pointer(dgk_last_error(), p, [512])
    do i = 1, 512
      if (p(i) == c_null_char) exit
      msg(i:i) = p(i)
    end do
  end function dgk_error_message

end module mod_dgk_hip
