!> Fortran host of the device advection path.
!!
!! Mirrors the reference driver (main.f90): the same namelist &PARAM_ADVECT3D
!! read from the file named by the first argument (main.f90:103-129), the same
!! initial condition sin(2 pi x) sin(2 pi y) sin(2 pi z) with constant velocity
!! (main.f90:146-157), the same SSP-RK3 loop and min/max trace format
!! (main.f90:51-80) and the same timer report (main.f90:175-177,
!! mod_advect3d_eq.f90:34-35).  The operator evaluation is selected by
!! DGOptrKernel_OptType: "HIP" (FP64) or "HIP_MP" (mixed precision); the CPU
!! kernels "GENERAL" and "OPT1" belong to the reference binary.
!!
!! Extra namelist entries (all optional, so reference input files still read):
!!   mesh_arrays   .true.  build the reference's per-node mesh arrays and hand
!!                         them to dgk_create (the drop-in route), .false. lets
!!                         the library build the structured mesh itself
!!   stage_loop    'FUSED' one fused kernel per RK stage (dgk_step), or
!!                 'PIECEWISE' the reference's stage loop call for call:
!!                 q0 copy, update_halo, cal_tend, RK update
!!   operator_dir  'operator_data'
!!   output_file   ''      when set, the final q(Np,Ne) is written there
!!                         (stream access, float64)
program dgk_main
  use, intrinsic :: iso_c_binding
  use, intrinsic :: iso_fortran_env, only: error_unit, real64, int64
  use mod_dgk
  implicit none

  integer, parameter :: RP = real64
  real(RP), parameter :: PI = 4.0_RP * atan(1.0_RP)

  integer :: NeX = 4, NeY = 4, NeZ = 4, PolyOrder = 3
  real(RP) :: vel_x = 1.0_RP, vel_y = 1.0_RP, vel_z = 1.0_RP
  character(len=8) :: DGOptrKernel_OptType = "HIP"
  real(RP) :: dt = 1.0e-3_RP
  integer :: nstep = 100, output_interval = 10
  logical :: mesh_arrays = .true.
  character(len=16) :: stage_loop = "FUSED"
  character(len=256) :: operator_dir = "operator_data"
  character(len=256) :: output_file = ""
  namelist /PARAM_ADVECT3D/ NeX, NeY, NeZ, PolyOrder, dt, nstep, output_interval, &
    vel_x, vel_y, vel_z, DGOptrKernel_OptType, mesh_arrays, stage_loop, operator_dir, &
    output_file

  character(len=256) :: conf_file
  integer :: fid, istep, stage
  integer(c_int) :: evaltype, Nq, Np, Nfp, NfpTot
  integer(c_int64_t) :: Ne, NeA, NhaloNode
  type(c_ptr) :: solver
  real(RP), allocatable :: x1D(:), D1D(:,:), Lift_mat(:,:,:,:)
  integer, allocatable :: VMapM(:,:), VMapP(:,:), halo_src_map(:)
  real(RP), allocatable :: normal_fn(:,:,:), Escale(:,:,:), Fscale(:,:), pos_en(:,:,:)
  real(RP), allocatable :: q(:,:), u(:,:), v(:,:), w(:,:)
  real(c_double) :: qmin, qmax, secs(5)
  integer(c_int64_t) :: counts(5)
  integer(int64) :: c0, c1, crate
  logical :: piecewise

  if (command_argument_count() < 1) then
    write(error_unit,'(A)') 'Error: configuration file argument is required.'
    write(error_unit,'(A)') 'Usage: dgk_main input.conf'
    flush(error_unit)
    error stop 1
  end if
  call get_command_argument(1, conf_file)
  open(newunit=fid, file=trim(conf_file), status='old', action='read')
  read(fid, nml=PARAM_ADVECT3D)
  close(fid)

  call dgk_check(dgk_evaltype_id(dgk_cstr(DGOptrKernel_OptType), evaltype), &
                 'DGOptrKernel_OptType')
  if (evaltype /= DGK_EVAL_HIP .and. evaltype /= DGK_EVAL_HIP_MP) then
    write(error_unit,'(A,A,A)') 'Error: DGOptrKernel_OptType = ', trim(DGOptrKernel_OptType), &
      ' names a reference CPU kernel; this host runs HIP or HIP_MP'
    error stop 1
  end if
  piecewise = trim(stage_loop) == 'PIECEWISE'

  !- mesh and operators (mesh_setup, mod_mesh.f90:74-193)
  call dgk_check(dgk_element_sizes(PolyOrder, Nq, Np, Nfp, NfpTot), 'element sizes')
  call dgk_check(dgk_mesh_sizes(NeX, NeY, NeZ, PolyOrder, Ne, NeA, NhaloNode), 'mesh sizes')
  allocate(x1D(Nq), D1D(Nq,Nq), Lift_mat(Nq,Nq,Nq,6))
  call dgk_check(dgk_load_operators(dgk_cstr(operator_dir), PolyOrder, x1D, D1D, Lift_mat), &
                 'operator data')
  allocate(VMapM(NfpTot,Ne), VMapP(NfpTot,Ne), halo_src_map(max(NhaloNode,1_int64)))
  allocate(normal_fn(NfpTot,Ne,3), Escale(Np,Ne,3), Fscale(NfpTot,Ne), pos_en(Np,Ne,3))
  call dgk_check(dgk_mesh_build(NeX, NeY, NeZ, PolyOrder, x1D, 1.0_RP, 1.0_RP, 1.0_RP, &
    VMapM, VMapP, halo_src_map, normal_fn, Escale, Fscale, pos_en), 'mesh build')

  if (mesh_arrays) then
    call dgk_check(dgk_create(evaltype, Nq, Ne, NeA, NhaloNode, D1D, Lift_mat, VMapM, VMapP, &
      halo_src_map, normal_fn, Escale, Fscale, solver), 'dgk_create')
  else
    call dgk_check(dgk_create_structured(evaltype, NeX, NeY, NeZ, PolyOrder, &
      dgk_cstr(operator_dir), 1.0_RP, 1.0_RP, 1.0_RP, solver), 'dgk_create_structured')
  end if

  !- initial condition and velocity (main.f90:146-157)
  allocate(q(Np,Ne), u(Np,Ne), v(Np,Ne), w(Np,Ne))
  q = sin(2.0_RP*PI*pos_en(:,:,1)) * sin(2.0_RP*PI*pos_en(:,:,2)) * sin(2.0_RP*PI*pos_en(:,:,3))
  u = vel_x
  v = vel_y
  w = vel_z
  call dgk_check(dgk_set_field(solver, DGK_FIELD_Q, q, int(Np,c_int64_t)*Ne), 'set q')
  call dgk_check(dgk_set_field(solver, DGK_FIELD_U, u, int(Np,c_int64_t)*Ne), 'set u')
  call dgk_check(dgk_set_field(solver, DGK_FIELD_V, v, int(Np,c_int64_t)*Ne), 'set v')
  call dgk_check(dgk_set_field(solver, DGK_FIELD_W, w, int(Np,c_int64_t)*Ne), 'set w')
  call dgk_check(dgk_set_kernel_timing(solver, 1_c_int), 'kernel timing')

  !- time loop (main.f90:51-80)
  call system_clock(c0, crate)
  do istep = 1, nstep
    if (piecewise) then
      call dgk_check(dgk_copy_q0(solver), 'copy q0')
      do stage = 1, 3
        call dgk_check(dgk_update_halo(solver, DGK_FIELD_Q), 'update_halo')
        call dgk_check(dgk_cal_tend(solver, c_null_ptr), 'cal_tend')
        call dgk_check(dgk_rk_update(solver, int(stage,c_int), dt), 'rk_update')
      end do
    else
      call dgk_check(dgk_step(solver, dt, 1_c_int), 'dgk_step')
    end if
    if (mod(istep, output_interval) == 0) then
      call dgk_check(dgk_minmax(solver, qmin, qmax), 'minmax')
      write(*,'(I8,2ES24.15)') istep, qmin, qmax
    end if
  end do
  call dgk_check(dgk_synchronize(solver), 'synchronize')
  call system_clock(c1)

  call dgk_check(dgk_timer_report(solver, secs, counts), 'timers')
  write(*,'(A)') "= Report of execution time [sec]"
  write(*,'(A30,ES24.5)') "Main:", real(c1 - c0, RP) / real(crate, RP)
  if (piecewise) then
    write(*,'(A30,ES24.5)') "Cal_tend:", secs(2)
  else
    write(*,'(A30,ES24.5)') "Cal_tend:", secs(5)
  end if
  write(*,'(A30,ES24.5)') "Element boundary flux:", secs(3)
  write(*,'(A30,ES24.5)') "Volume derivate + surface lift:", secs(4)

  if (len_trim(output_file) > 0) then
    call dgk_check(dgk_get_field(solver, DGK_FIELD_Q, q, int(Np,c_int64_t)*Ne), 'get q')
    open(newunit=fid, file=trim(output_file), access='stream', form='unformatted', &
         status='replace', action='write')
    write(fid) q
    close(fid)
  end if
  call dgk_check(dgk_destroy(solver), 'destroy')
end program dgk_main
