!-------------------------------------------------------------------------------
!> ref_harness: a driver that links the reference's own modules (compiled from
!! source under the reference tree by oracle/Makefile into oracle/_ref/) and
!! exposes them for golden-vector generation and CPU-baseline timing.
!!
!! TEST INFRASTRUCTURE ONLY: it replaces the reference's main.f90 because that
!! program hard-codes its initial condition (main.f90:146-157) and prints only
!! min/max.  The step loop below performs the same calls in the same order as
!! main.f90:51-80 (q0 <- q; per stage update_halo, advect3d_eq_cal_tend, RK
!! combine) so that the trajectories are the reference's.
!!
!! Usage (all arrays are raw little-endian float64 / int32, Fortran order):
!!   ref_harness mesh <NeX> <NeY> <NeZ> <p> <outfile>
!!   ref_harness optr <p> <GENERAL|OPT1> <nelem> <infile> <outfile>
!!   ref_harness optr2 <p> <nelem> <infile> <outfile>
!!   ref_harness tend <conf> <icfile> <outfile>
!!   ref_harness run  <conf> <icfile> <outfile>
!! The working directory must hold operator_data/p<p>.dat (mod_mesh.f90:253).
!-------------------------------------------------------------------------------
program ref_harness
  use mod_common, only: RP, rk_a => RK3s3oSSP_rk_a, rk_b => RK3s3oSSP_rk_b, &
       RK_nstage => RK3s3oSSP_nstage
  use mod_mesh, only: mesh_setup, update_halo, Nq, Np, Nfp, NfpTot, Ne, NeA, &
       x1D, D1D, D1D_tr, Lift_mat, VMapM, VMapP, normal_fn, Escale, Fscale, pos_en
  use mod_dg_optr_kernel, only: dg_optr_kernel_setup, &
       tensorprod_divlike_dirXYZ, tensorprod_Lift_hexahedral
  use mod_advect3d_eq, only: advect3d_eq_cal_tend
  use mod_dg_optr_kernel_opt1
  use omp_lib, only: omp_get_wtime, omp_get_max_threads
  implicit none

  character(len=256) :: mode, a1, a2, a3, a4, a5
  integer :: nx, ny, nz, p, nelem

  call get_command_argument(1, mode)
  select case (trim(mode))
  case ("mesh")
    call get_command_argument(2, a1); read(a1,*) nx
    call get_command_argument(3, a2); read(a2,*) ny
    call get_command_argument(4, a3); read(a3,*) nz
    call get_command_argument(5, a4); read(a4,*) p
    call get_command_argument(6, a5)
    call do_mesh(nx, ny, nz, p, trim(a5))
  case ("optr")
    call get_command_argument(2, a1); read(a1,*) p
    call get_command_argument(3, a2)
    call get_command_argument(4, a3); read(a3,*) nelem
    call get_command_argument(5, a4)
    call get_command_argument(6, a5)
    call do_optr(p, trim(a2), nelem, trim(a4), trim(a5))
  case ("optr2")
    call get_command_argument(2, a1); read(a1,*) p
    call get_command_argument(3, a3); read(a3,*) nelem
    call get_command_argument(4, a4)
    call get_command_argument(5, a5)
    call do_optr2(p, nelem, trim(a4), trim(a5))
  case ("tend", "run")
    call get_command_argument(2, a1)
    call get_command_argument(3, a2)
    call get_command_argument(4, a3)
    call do_run(trim(mode), trim(a1), trim(a2), trim(a3))
  case ("bench")
    call get_command_argument(2, a1)
    call get_command_argument(3, a2)
    call get_command_argument(4, a3)
    call get_command_argument(5, a4); read(a4,*) nelem
    call do_bench(trim(a1), trim(a2), trim(a3), nelem)
  case default
    write(*,'(A)') "usage: ref_harness mesh|optr|tend|run|bench ..."
    error stop 2
  end select

contains

  subroutine do_mesh(nx, ny, nz, p, outfile)
    integer, intent(in) :: nx, ny, nz, p
    character(len=*), intent(in) :: outfile
    real(RP), allocatable :: f(:)
    integer :: nhn, ib, fid
    call mesh_setup(nx, ny, nz, p, 1.0_RP, 1.0_RP, 1.0_RP)
    ! halo_src_map is private to mod_mesh: recover it through update_halo on
    ! a field that holds its own 1-based flat index.
    nhn = Nfp * 2 * (ny*nz + nx*nz + nx*ny)
    allocate(f(Np*NeA))
    do ib = 1, Np*NeA
      f(ib) = real(ib, RP)
    end do
    call update_halo(f)
    open(newunit=fid, file=outfile, access='stream', form='unformatted', status='replace')
    write(fid) Nq, Np, NfpTot, Ne, NeA, nhn
    write(fid) x1D, D1D, D1D_tr, Lift_mat
    write(fid) VMapM, VMapP
    write(fid) int(f(Np*Ne+1:Np*Ne+nhn))
    write(fid) normal_fn, Escale, Fscale, pos_en
    close(fid)
  end subroutine do_mesh

  subroutine do_optr(p, tname, nelem, infile, outfile)
    integer, intent(in) :: p, nelem
    character(len=*), intent(in) :: tname, infile, outfile
    real(RP), allocatable :: ix(:,:), iy(:,:), iz(:,:), fb(:,:)
    real(RP), allocatable :: ox(:,:), oy(:,:), oz(:,:), ol(:,:)
    integer :: fid, ke
    call mesh_setup(1, 1, 1, p, 1.0_RP, 1.0_RP, 1.0_RP)
    call dg_optr_kernel_setup(tname)
    allocate(ix(Np,nelem), iy(Np,nelem), iz(Np,nelem), fb(NfpTot,nelem))
    allocate(ox(Np,nelem), oy(Np,nelem), oz(Np,nelem), ol(Np,nelem))
    open(newunit=fid, file=infile, access='stream', form='unformatted', status='old')
    read(fid) ix, iy, iz, fb
    close(fid)
    do ke = 1, nelem
      call tensorprod_divlike_dirXYZ(ox(:,ke), oy(:,ke), oz(:,ke), D1D, D1D_tr, &
           ix(:,ke), iy(:,ke), iz(:,ke), Nq)
      call tensorprod_Lift_hexahedral(ol(:,ke), Lift_mat, fb(:,ke), Nq)
    end do
    open(newunit=fid, file=outfile, access='stream', form='unformatted', status='replace')
    write(fid) ox, oy, oz, ol
    close(fid)
  end subroutine do_optr

  !> optr2: the OPT1 module's directional kernels dirX/dirY/dirZ
  !! (mod_dg_optr_kernel_opt1.F90.erb:51-129) and gradlike_dirXYZ (:166-219)
  !! over nelem element blocks a(Np,nelem), with D1D (x) and D1D_tr (y, z).
  subroutine do_optr2(p, nelem, infile, outfile)
    integer, intent(in) :: p, nelem
    character(len=*), intent(in) :: infile, outfile
    real(RP), allocatable :: a(:,:), ox(:,:), oy(:,:), oz(:,:), gx(:,:), gy(:,:), gz(:,:)
    integer :: fid, ke
    call mesh_setup(1, 1, 1, p, 1.0_RP, 1.0_RP, 1.0_RP)
    allocate(a(Np,nelem), ox(Np,nelem), oy(Np,nelem), oz(Np,nelem))
    allocate(gx(Np,nelem), gy(Np,nelem), gz(Np,nelem))
    open(newunit=fid, file=infile, access='stream', form='unformatted', status='old')
    read(fid) a
    close(fid)
    do ke = 1, nelem
      select case (p)
      case (1)
        call element_operation_kernel_matvec_dirX_P1(D1D, a(:,ke), ox(:,ke))
        call element_operation_kernel_matvec_dirY_P1(D1D_tr, a(:,ke), oy(:,ke))
        call element_operation_kernel_matvec_dirZ_P1(D1D_tr, a(:,ke), oz(:,ke))
        call element_operation_kernel_matvec_gradlike_dirXYZ_P1(D1D, D1D_tr, a(:,ke), a(:,ke), &
             gx(:,ke), gy(:,ke), gz(:,ke))
      case (2)
        call element_operation_kernel_matvec_dirX_P2(D1D, a(:,ke), ox(:,ke))
        call element_operation_kernel_matvec_dirY_P2(D1D_tr, a(:,ke), oy(:,ke))
        call element_operation_kernel_matvec_dirZ_P2(D1D_tr, a(:,ke), oz(:,ke))
        call element_operation_kernel_matvec_gradlike_dirXYZ_P2(D1D, D1D_tr, a(:,ke), a(:,ke), &
             gx(:,ke), gy(:,ke), gz(:,ke))
      case (3)
        call element_operation_kernel_matvec_dirX_P3(D1D, a(:,ke), ox(:,ke))
        call element_operation_kernel_matvec_dirY_P3(D1D_tr, a(:,ke), oy(:,ke))
        call element_operation_kernel_matvec_dirZ_P3(D1D_tr, a(:,ke), oz(:,ke))
        call element_operation_kernel_matvec_gradlike_dirXYZ_P3(D1D, D1D_tr, a(:,ke), a(:,ke), &
             gx(:,ke), gy(:,ke), gz(:,ke))
      case (4)
        call element_operation_kernel_matvec_dirX_P4(D1D, a(:,ke), ox(:,ke))
        call element_operation_kernel_matvec_dirY_P4(D1D_tr, a(:,ke), oy(:,ke))
        call element_operation_kernel_matvec_dirZ_P4(D1D_tr, a(:,ke), oz(:,ke))
        call element_operation_kernel_matvec_gradlike_dirXYZ_P4(D1D, D1D_tr, a(:,ke), a(:,ke), &
             gx(:,ke), gy(:,ke), gz(:,ke))
      case (5)
        call element_operation_kernel_matvec_dirX_P5(D1D, a(:,ke), ox(:,ke))
        call element_operation_kernel_matvec_dirY_P5(D1D_tr, a(:,ke), oy(:,ke))
        call element_operation_kernel_matvec_dirZ_P5(D1D_tr, a(:,ke), oz(:,ke))
        call element_operation_kernel_matvec_gradlike_dirXYZ_P5(D1D, D1D_tr, a(:,ke), a(:,ke), &
             gx(:,ke), gy(:,ke), gz(:,ke))
      case (6)
        call element_operation_kernel_matvec_dirX_P6(D1D, a(:,ke), ox(:,ke))
        call element_operation_kernel_matvec_dirY_P6(D1D_tr, a(:,ke), oy(:,ke))
        call element_operation_kernel_matvec_dirZ_P6(D1D_tr, a(:,ke), oz(:,ke))
        call element_operation_kernel_matvec_gradlike_dirXYZ_P6(D1D, D1D_tr, a(:,ke), a(:,ke), &
             gx(:,ke), gy(:,ke), gz(:,ke))
      case (7)
        call element_operation_kernel_matvec_dirX_P7(D1D, a(:,ke), ox(:,ke))
        call element_operation_kernel_matvec_dirY_P7(D1D_tr, a(:,ke), oy(:,ke))
        call element_operation_kernel_matvec_dirZ_P7(D1D_tr, a(:,ke), oz(:,ke))
        call element_operation_kernel_matvec_gradlike_dirXYZ_P7(D1D, D1D_tr, a(:,ke), a(:,ke), &
             gx(:,ke), gy(:,ke), gz(:,ke))
      case (8)
        call element_operation_kernel_matvec_dirX_P8(D1D, a(:,ke), ox(:,ke))
        call element_operation_kernel_matvec_dirY_P8(D1D_tr, a(:,ke), oy(:,ke))
        call element_operation_kernel_matvec_dirZ_P8(D1D_tr, a(:,ke), oz(:,ke))
        call element_operation_kernel_matvec_gradlike_dirXYZ_P8(D1D, D1D_tr, a(:,ke), a(:,ke), &
             gx(:,ke), gy(:,ke), gz(:,ke))
      case (9)
        call element_operation_kernel_matvec_dirX_P9(D1D, a(:,ke), ox(:,ke))
        call element_operation_kernel_matvec_dirY_P9(D1D_tr, a(:,ke), oy(:,ke))
        call element_operation_kernel_matvec_dirZ_P9(D1D_tr, a(:,ke), oz(:,ke))
        call element_operation_kernel_matvec_gradlike_dirXYZ_P9(D1D, D1D_tr, a(:,ke), a(:,ke), &
             gx(:,ke), gy(:,ke), gz(:,ke))
      case (10)
        call element_operation_kernel_matvec_dirX_P10(D1D, a(:,ke), ox(:,ke))
        call element_operation_kernel_matvec_dirY_P10(D1D_tr, a(:,ke), oy(:,ke))
        call element_operation_kernel_matvec_dirZ_P10(D1D_tr, a(:,ke), oz(:,ke))
        call element_operation_kernel_matvec_gradlike_dirXYZ_P10(D1D, D1D_tr, a(:,ke), a(:,ke), &
             gx(:,ke), gy(:,ke), gz(:,ke))
      case (11)
        call element_operation_kernel_matvec_dirX_P11(D1D, a(:,ke), ox(:,ke))
        call element_operation_kernel_matvec_dirY_P11(D1D_tr, a(:,ke), oy(:,ke))
        call element_operation_kernel_matvec_dirZ_P11(D1D_tr, a(:,ke), oz(:,ke))
        call element_operation_kernel_matvec_gradlike_dirXYZ_P11(D1D, D1D_tr, a(:,ke), a(:,ke), &
             gx(:,ke), gy(:,ke), gz(:,ke))
      case (12)
        call element_operation_kernel_matvec_dirX_P12(D1D, a(:,ke), ox(:,ke))
        call element_operation_kernel_matvec_dirY_P12(D1D_tr, a(:,ke), oy(:,ke))
        call element_operation_kernel_matvec_dirZ_P12(D1D_tr, a(:,ke), oz(:,ke))
        call element_operation_kernel_matvec_gradlike_dirXYZ_P12(D1D, D1D_tr, a(:,ke), a(:,ke), &
             gx(:,ke), gy(:,ke), gz(:,ke))
      case (13)
        call element_operation_kernel_matvec_dirX_P13(D1D, a(:,ke), ox(:,ke))
        call element_operation_kernel_matvec_dirY_P13(D1D_tr, a(:,ke), oy(:,ke))
        call element_operation_kernel_matvec_dirZ_P13(D1D_tr, a(:,ke), oz(:,ke))
        call element_operation_kernel_matvec_gradlike_dirXYZ_P13(D1D, D1D_tr, a(:,ke), a(:,ke), &
             gx(:,ke), gy(:,ke), gz(:,ke))
      case (14)
        call element_operation_kernel_matvec_dirX_P14(D1D, a(:,ke), ox(:,ke))
        call element_operation_kernel_matvec_dirY_P14(D1D_tr, a(:,ke), oy(:,ke))
        call element_operation_kernel_matvec_dirZ_P14(D1D_tr, a(:,ke), oz(:,ke))
        call element_operation_kernel_matvec_gradlike_dirXYZ_P14(D1D, D1D_tr, a(:,ke), a(:,ke), &
             gx(:,ke), gy(:,ke), gz(:,ke))
      case (15)
        call element_operation_kernel_matvec_dirX_P15(D1D, a(:,ke), ox(:,ke))
        call element_operation_kernel_matvec_dirY_P15(D1D_tr, a(:,ke), oy(:,ke))
        call element_operation_kernel_matvec_dirZ_P15(D1D_tr, a(:,ke), oz(:,ke))
        call element_operation_kernel_matvec_gradlike_dirXYZ_P15(D1D, D1D_tr, a(:,ke), a(:,ke), &
             gx(:,ke), gy(:,ke), gz(:,ke))
      end select
    end do
    open(newunit=fid, file=outfile, access='stream', form='unformatted', status='replace')
    write(fid) ox, oy, oz, gx, gy, gz
    close(fid)
  end subroutine do_optr2

  !> tend: one advect3d_eq_cal_tend on the initial state (after halo update).
  !! run : the main.f90 step loop; writes the final owned q, then per-step
  !!       (min,max) of the owned q, then the loop wall time and thread count.
  subroutine do_run(mode, conf, icfile, outfile)
    character(len=*), intent(in) :: mode, conf, icfile, outfile
    integer :: NeX = 4, NeY = 4, NeZ = 4, PolyOrder = 3
    integer :: nstep = 1, output_interval = 1
    real(RP) :: dt = 1.0e-3_RP
    real(RP) :: vel_x = 1.0_RP, vel_y = 1.0_RP, vel_z = 1.0_RP
    character(len=8) :: DGOptrKernel_OptType = "OPT1"
    namelist /PARAM_ADVECT3D/ NeX, NeY, NeZ, PolyOrder, dt, nstep, &
         output_interval, vel_x, vel_y, vel_z, DGOptrKernel_OptType
    real(RP), allocatable :: q(:,:), q0(:,:), dqdt(:,:), u(:,:), v(:,:), w(:,:)
    real(RP), allocatable :: mm(:,:)
    integer :: fid, istep, stage, ke
    real(RP) :: t0, t1

    open(newunit=fid, file=conf, status='old', action='read')
    read(fid, nml=PARAM_ADVECT3D)
    close(fid)
    call mesh_setup(NeX, NeY, NeZ, PolyOrder, 1.0_RP, 1.0_RP, 1.0_RP)
    allocate(q(Np,NeA), q0(Np,NeA), dqdt(Np,NeA))
    allocate(u(Np,NeA), v(Np,NeA), w(Np,NeA), mm(2,max(nstep,1)))
    call dg_optr_kernel_setup(DGOptrKernel_OptType)
    q = 0.0_RP; u = 0.0_RP; v = 0.0_RP; w = 0.0_RP; q0 = 0.0_RP; dqdt = 0.0_RP
    open(newunit=fid, file=icfile, access='stream', form='unformatted', status='old')
    read(fid) q(:,1:Ne), u(:,1:Ne), v(:,1:Ne), w(:,1:Ne)
    close(fid)
    call update_halo(q)
    call update_halo(u)
    call update_halo(v)
    call update_halo(w)

    if (mode == "tend") then
      call advect3d_eq_cal_tend(dqdt, q, u, v, w, D1D, D1D_tr, Lift_mat, &
           VMapM, VMapP, normal_fn, Escale, Fscale, Nq, Np, NfpTot, Ne, NeA)
      open(newunit=fid, file=outfile, access='stream', form='unformatted', status='replace')
      write(fid) dqdt(:,1:Ne)
      close(fid)
      return
    end if

    t0 = omp_get_wtime()
    do istep = 1, nstep
      !$omp parallel do
      do ke = 1, Ne
        q0(:,ke) = q(:,ke)
      end do
      do stage = 1, RK_nstage
        call update_halo(q)
        call advect3d_eq_cal_tend(dqdt, q, u, v, w, D1D, D1D_tr, Lift_mat, &
             VMapM, VMapP, normal_fn, Escale, Fscale, Nq, Np, NfpTot, Ne, NeA)
        !$omp parallel do
        do ke = 1, Ne
          q(:,ke) = rk_a(stage) * q0(:,ke) &
               + rk_b(stage) * ( q(:,ke) + dt * dqdt(:,ke) )
        end do
      end do
      mm(1,istep) = minval(q(:,1:Ne))
      mm(2,istep) = maxval(q(:,1:Ne))
      if (mod(istep, output_interval) == 0) then
        write(*,'(I8,2ES24.15)') istep, mm(1,istep), mm(2,istep)
      end if
    end do
    t1 = omp_get_wtime()
    write(*,'(A,ES24.15,A,I6)') "LOOP_SECONDS ", t1 - t0, " THREADS ", omp_get_max_threads()
    open(newunit=fid, file=outfile, access='stream', form='unformatted', status='replace')
    write(fid) q(:,1:Ne)
    write(fid) mm(:,1:nstep)
    write(fid) t1 - t0
    close(fid)
  end subroutine do_run

  !> bench: the main.f90 step loop timed nrep times from the same initial
  !! state with each operator type, GENERAL then OPT1 (one mesh setup and one
  !! read of the initial fields for all of them); writes the 2*nrep loop
  !! times (GENERAL's, then OPT1's) and the thread count.
  subroutine do_bench(conf, icfile, outfile, nrep)
    character(len=*), intent(in) :: conf, icfile, outfile
    integer, intent(in) :: nrep
    integer :: NeX = 4, NeY = 4, NeZ = 4, PolyOrder = 3
    integer :: nstep = 1, output_interval = 1
    real(RP) :: dt = 1.0e-3_RP
    real(RP) :: vel_x = 1.0_RP, vel_y = 1.0_RP, vel_z = 1.0_RP
    character(len=8) :: DGOptrKernel_OptType = "OPT1"
    namelist /PARAM_ADVECT3D/ NeX, NeY, NeZ, PolyOrder, dt, nstep, &
         output_interval, vel_x, vel_y, vel_z, DGOptrKernel_OptType
    real(RP), allocatable :: q(:,:), qi(:,:), q0(:,:), dqdt(:,:), u(:,:), v(:,:), w(:,:)
    real(RP), allocatable :: secs(:)
    character(len=8), parameter :: types(2) = [character(len=8) :: "GENERAL", "OPT1"]
    integer :: fid, istep, stage, ke, it, r
    real(RP) :: t0

    open(newunit=fid, file=conf, status='old', action='read')
    read(fid, nml=PARAM_ADVECT3D)
    close(fid)
    call mesh_setup(NeX, NeY, NeZ, PolyOrder, 1.0_RP, 1.0_RP, 1.0_RP)
    allocate(q(Np,NeA), qi(Np,Ne), q0(Np,NeA), dqdt(Np,NeA))
    allocate(u(Np,NeA), v(Np,NeA), w(Np,NeA), secs(2*nrep))
    q = 0.0_RP; u = 0.0_RP; v = 0.0_RP; w = 0.0_RP; q0 = 0.0_RP; dqdt = 0.0_RP
    open(newunit=fid, file=icfile, access='stream', form='unformatted', status='old')
    read(fid) qi(:,1:Ne), u(:,1:Ne), v(:,1:Ne), w(:,1:Ne)
    close(fid)
    call update_halo(u)
    call update_halo(v)
    call update_halo(w)
    do it = 1, 2
      call dg_optr_kernel_setup(types(it))
      do r = 1, nrep
        !$omp parallel do
        do ke = 1, Ne
          q(:,ke) = qi(:,ke)
        end do
        call update_halo(q)
        t0 = omp_get_wtime()
        do istep = 1, nstep
          !$omp parallel do
          do ke = 1, Ne
            q0(:,ke) = q(:,ke)
          end do
          do stage = 1, RK_nstage
            call update_halo(q)
            call advect3d_eq_cal_tend(dqdt, q, u, v, w, D1D, D1D_tr, Lift_mat, &
                 VMapM, VMapP, normal_fn, Escale, Fscale, Nq, Np, NfpTot, Ne, NeA)
            !$omp parallel do
            do ke = 1, Ne
              q(:,ke) = rk_a(stage) * q0(:,ke) &
                   + rk_b(stage) * ( q(:,ke) + dt * dqdt(:,ke) )
            end do
          end do
        end do
        secs((it-1)*nrep + r) = omp_get_wtime() - t0
        write(*,'(A,A8,A,I3,A,ES16.6)') "BENCH ", types(it), " rep", r, " seconds", &
             secs((it-1)*nrep + r)
      end do
    end do
    open(newunit=fid, file=outfile, access='stream', form='unformatted', status='replace')
    write(fid) secs
    write(fid) real(omp_get_max_threads(), RP)
    close(fid)
  end subroutine do_bench

end program ref_harness
