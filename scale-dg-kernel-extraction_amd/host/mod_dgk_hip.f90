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
  integer, public, parameter :: DGK_TYPE_GENERAL = 1, DGK_TYPE_OPT1 = 2, DGK_TYPE_HIP = 3
  integer, public, parameter :: DGK_FLAG_FMA = 1, DGK_FLAG_FORCE_TABLES = 2, &
       DGK_FLAG_NODE_KERNEL = 4, DGK_FLAG_FACTORED_LIFT = 8
  integer, public, parameter :: DGK_FIELD_Q = 0, DGK_FIELD_Q0 = 1, DGK_FIELD_DQDT = 2

  public :: dgk_optr_kernel_setup, dgk_read_operator, dgk_mesh_sizes, dgk_mesh_setup
  public :: dgk_create, dgk_create_structured, dgk_destroy
  public :: dgk_set_fields, dgk_get_field, dgk_cal_tend, dgk_rk_stage, dgk_step, dgk_run
  public :: dgk_synchronize, dgk_advect3d_eq_cal_tend, dgk_error_message

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

    type(c_ptr) function dgk_last_error() bind(C, name="dgk_last_error")
      import :: c_ptr
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
