!-------------------------------------------------------------------------------
!> dg_advect3d: Fortran host of the 3D DG advection run on the MI355X.
!!
!! Reads the reference's configuration namelist &PARAM_ADVECT3D (same knobs
!! and defaults as main.f90:94-115: NeX, NeY, NeZ, PolyOrder, dt, nstep,
!! output_interval, vel_x/y/z, DGOptrKernel_OptType) and runs the SSP-RK3
!! time loop of main.f90:51-80 on the device through mod_dgk_hip.  The
!! initial field is the reference's q = sin(2 pi x) sin(2 pi y) sin(2 pi z)
!! with constant velocity (main.f90:146-157); the diagnostics line every
!! output_interval steps is the reference's "(I8,2ES24.15) istep, min, max"
!! (main.f90:76-79).
!!
!! Knobs of this host only are in a separate group, so that a reference
!! input.conf is read unchanged:
!!   &PARAM_DGK_HIP operator_data_dir, connectivity ("STRUCTURED" builds the
!!     compressed mesh on the device side, "TABLES" passes the reference's
!!     per-node tables as the reference's mesh_setup produces them),
!!     kernel_flags (DGK_FLAG_*), sum_order ("OPT1" or "GENERAL": the
!!     reference kernels' summation order the device reproduces),
!!     dump_initial / dump_final (raw float64 files of the owned q(Np,Ne)) /
!!
!! Usage: dg_advect3d input.conf
!-------------------------------------------------------------------------------
program dg_advect3d
  use, intrinsic :: iso_c_binding
  use, intrinsic :: iso_fortran_env, only: error_unit, int64
  use mod_dgk_hip
  implicit none

  integer :: NeX = 4, NeY = 4, NeZ = 4, PolyOrder = 3
  integer :: nstep = 100, output_interval = 10
  real(c_double) :: dt = 1.0e-3_c_double
  real(c_double) :: vel_x = 1.0_c_double, vel_y = 1.0_c_double, vel_z = 1.0_c_double
  character(len=8) :: DGOptrKernel_OptType = "HIP"
  namelist /PARAM_ADVECT3D/ NeX, NeY, NeZ, PolyOrder, dt, nstep, output_interval, &
       vel_x, vel_y, vel_z, DGOptrKernel_OptType

  character(len=256) :: operator_data_dir = "operator_data"
  character(len=16) :: connectivity = "STRUCTURED"
  character(len=256) :: dump_initial = "", dump_final = ""
  character(len=8) :: sum_order = "OPT1"
  integer :: kernel_flags = 0
  namelist /PARAM_DGK_HIP/ operator_data_dir, connectivity, kernel_flags, sum_order, &
       dump_initial, dump_final

  real(c_double), parameter :: PI = 4.0_c_double * atan(1.0_c_double)
  character(len=256) :: conf
  integer :: fid, ios, type_id, order_id, rc, sizes(7)
  integer :: Nq, Np, NfpTot, Ne, NeA, NhaloNode, ke, n, istep
  real(c_double), allocatable :: x1D(:), D1D(:,:), Lift(:,:,:,:)
  real(c_double), allocatable :: q(:,:), u(:,:), v(:,:), w(:,:), pos_en(:,:,:), mm(:,:)
  integer(c_int), allocatable :: VMapM(:,:), VMapP(:,:), halo_src_map(:)
  real(c_double), allocatable :: normal_fn(:,:,:), Escale(:,:,:), Fscale(:,:)
  type(c_ptr) :: solver
  integer(int64) :: c0, c1, crate

  if (command_argument_count() < 1) then
    write(error_unit,'(A)') 'Error: configuration file argument is required.'
    write(error_unit,'(A)') 'Usage: dg_advect3d input.conf'
    error stop 1
  end if
  call get_command_argument(1, conf)
  open(newunit=fid, file=trim(conf), status='old', action='read')
  read(fid, nml=PARAM_ADVECT3D)
  rewind(fid)
  read(fid, nml=PARAM_DGK_HIP, iostat=ios)  ! optional group
  close(fid)

  rc = dgk_optr_kernel_setup(trim(DGOptrKernel_OptType)//c_null_char, type_id)
  call check(rc, "dgk_optr_kernel_setup")
  if (type_id /= DGK_TYPE_HIP) then
    write(error_unit,'(A)') 'dg_advect3d runs DGOptrKernel_OptType = "HIP"; GENERAL and OPT1 '// &
         'are the reference''s CPU kernels'
    error stop 2
  end if

  rc = dgk_mesh_sizes(NeX, NeY, NeZ, PolyOrder, sizes)
  call check(rc, "dgk_mesh_sizes")
  Nq = sizes(1); Np = sizes(2); NfpTot = sizes(4); Ne = sizes(5); NeA = sizes(6)
  NhaloNode = sizes(7)
  allocate(x1D(Nq), D1D(Nq,Nq), Lift(Nq,Nq,Nq,6))
  rc = dgk_read_operator(trim(operator_data_dir)//c_null_char, PolyOrder, x1D, D1D, Lift)
  call check(rc, "dgk_read_operator")

  allocate(q(Np,NeA), u(Np,NeA), v(Np,NeA), w(Np,NeA), pos_en(Np,Ne,3))
  allocate(VMapM(NfpTot,Ne), VMapP(NfpTot,Ne), halo_src_map(max(NhaloNode,1)))
  allocate(normal_fn(NfpTot,Ne,3), Escale(Np,Ne,3), Fscale(NfpTot,Ne))
  rc = dgk_mesh_setup(NeX, NeY, NeZ, PolyOrder, 1.0_c_double, 1.0_c_double, 1.0_c_double, x1D, &
       VMapM, VMapP, normal_fn, Escale, Fscale, pos_en, halo_src_map)
  call check(rc, "dgk_mesh_setup")

  q = 0.0_c_double; u = 0.0_c_double; v = 0.0_c_double; w = 0.0_c_double
  do ke = 1, Ne
    do n = 1, Np
      q(n,ke) = sin(2.0_c_double*PI*pos_en(n,ke,1)) * sin(2.0_c_double*PI*pos_en(n,ke,2)) &
              * sin(2.0_c_double*PI*pos_en(n,ke,3))
      u(n,ke) = vel_x
      v(n,ke) = vel_y
      w(n,ke) = vel_z
    end do
  end do
  if (len_trim(dump_initial) > 0) call dump(trim(dump_initial), q(:,1:Ne))

  ! the device evaluates in GENERAL's or OPT1's summation order
  rc = dgk_optr_kernel_setup(trim(sum_order)//c_null_char, order_id)
  call check(rc, "sum_order")
  select case (trim(connectivity))
  case ("TABLES")
    rc = dgk_create(solver, Nq, Ne, NeA, NfpTot, D1D, Lift, VMapM, VMapP, normal_fn, Escale, &
         Fscale, halo_src_map, NhaloNode, order_id, kernel_flags)
  case default
    rc = dgk_create_structured(solver, NeX, NeY, NeZ, PolyOrder, 1.0_c_double, 1.0_c_double, &
         1.0_c_double, D1D, Lift, order_id, kernel_flags)
  end select
  call check(rc, "dgk_create")
  rc = dgk_set_fields(solver, q, u, v, w)
  call check(rc, "dgk_set_fields")

  allocate(mm(2, max(nstep,1)))
  call system_clock(c0, crate)
  rc = dgk_run(solver, dt, nstep, mm)
  call check(rc, "dgk_run")
  call system_clock(c1)
  do istep = 1, nstep
    if (mod(istep, output_interval) == 0) write(*,'(I8,2ES24.15)') istep, mm(1,istep), mm(2,istep)
  end do

  rc = dgk_get_field(solver, DGK_FIELD_Q, q)
  call check(rc, "dgk_get_field")
  if (len_trim(dump_final) > 0) call dump(trim(dump_final), q(:,1:Ne))
  write(*,'(A)') "= Report of execution time [sec]"
  write(*,'(A30,ES24.5)') "Main:", real(c1 - c0, c_double) / real(crate, c_double)
  write(*,'(A30,ES24.5)') "GDOF-updates/s per RK stage:", &
       real(Np, c_double) * real(Ne, c_double) * 3.0_c_double * real(nstep, c_double) &
       / max(real(c1 - c0, c_double) / real(crate, c_double), 1.0e-30_c_double) / 1.0e9_c_double
  rc = dgk_destroy(solver)

contains

  subroutine check(code, what)
    integer, intent(in) :: code
    character(len=*), intent(in) :: what
    if (code /= DGK_OK) then
      write(error_unit,'(A,A,I0,A,A)') what, " failed (", code, "): ", trim(dgk_error_message())
      error stop 3
    end if
  end subroutine check

  subroutine dump(path, a)
    character(len=*), intent(in) :: path
    real(c_double), intent(in) :: a(:,:)
    integer :: f
    open(newunit=f, file=path, access='stream', form='unformatted', status='replace')
    write(f) a
    close(f)
  end subroutine dump

end program dg_advect3d
