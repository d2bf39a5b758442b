!> The reference's own mesh arrays through libdgk (TEST INFRASTRUCTURE ONLY).
!!
!! Builds the mesh with the reference's mesh_setup (mod_mesh.f90:74-193),
!! compiled by oracle/Makefile from the reference sources, and hands its
!! arrays -- D1D, Lift_mat, VMapM, VMapP, normal_fn, Escale, Fscale and the
!! halo map -- unchanged to dgk_create through the committed iso_c_binding
!! module (fortran/mod_dgk.f90), which is the route that INTEGRATION.md adds
!! to the reference's main.f90.  The device then runs nstep SSP-RK3 steps
!! (dgk_step), and the same process runs the reference's own modules on the
!! CPU over the same inputs (the stage loop of main.f90:51-75 with the
!! requested DGOptrKernel_OptType), so both traces and the final fields can
!! be compared.
!!
!! The halo map is private to mod_mesh.  It is read back through the public
!! update_halo (mod_mesh.f90:209-221) applied to a field that holds each
!! owned node's own flat index: afterwards f(Np*Ne+ib) = halo_src_map(ib).
!!
!! Usage (cwd must contain operator_data/):
!!   ref_dgk_host NeX NeY NeZ p DEVTYPE CPUTYPE nstep dt
!!   DEVTYPE = HIP | HIP_MP, CPUTYPE = GENERAL | OPT1
!! Initial condition and velocity as the reference driver (main.f90:146-157):
!! q = sin(2 pi x) sin(2 pi y) sin(2 pi z), (u, v, w) = (1.0, 0.5, 0.25).
!! stdout: per step "STEP istep dev_min dev_max cpu_min cpu_max", then
!! "MAXREL r" (max-norm relative difference of the final q) and "RELL2 r".
program ref_dgk_host
  use, intrinsic :: iso_c_binding
  use mod_common, only: RP, PI, &
    RK_nstage => RK3s3oSSP_nstage, rk_a => RK3s3oSSP_rk_a, rk_b => RK3s3oSSP_rk_b
  use mod_mesh, only: mesh_setup, update_halo, &
    Nq, Np, Nfp, NfpTot, Ne, NeA, D1D, D1D_tr, Lift_mat, VMapM, VMapP, &
    normal_fn, Escale, Fscale, pos_en
  use mod_advect3d_eq, only: advect3d_eq_cal_tend
  use mod_dg_optr_kernel, only: dg_optr_kernel_setup
  use mod_dgk
  implicit none

  character(len=256) :: arg, devtype, cputype
  integer :: NeX, NeY, NeZ, p, nstep, istep, stage, ke, n, ib, nhalo_node
  real(RP) :: dt
  real(RP), allocatable :: q(:,:), q0(:,:), dqdt(:,:), u(:,:), v(:,:), w(:,:)
  real(RP), allocatable :: idx(:,:), qdev(:,:)
  integer, allocatable :: halo_src_map(:)
  type(c_ptr) :: solver
  integer(c_int) :: evaltype
  integer(c_int64_t) :: nown
  real(c_double) :: dmin, dmax

  if (command_argument_count() < 8) then
    write(*,'(A)') 'usage: ref_dgk_host NeX NeY NeZ p DEVTYPE CPUTYPE nstep dt'
    error stop 1
  end if
  call get_command_argument(1, arg); read(arg,*) NeX
  call get_command_argument(2, arg); read(arg,*) NeY
  call get_command_argument(3, arg); read(arg,*) NeZ
  call get_command_argument(4, arg); read(arg,*) p
  call get_command_argument(5, devtype)
  call get_command_argument(6, cputype)
  call get_command_argument(7, arg); read(arg,*) nstep
  call get_command_argument(8, arg); read(arg,*) dt

  call mesh_setup(NeX, NeY, NeZ, p, 1.0_RP, 1.0_RP, 1.0_RP)
  allocate(q(Np,NeA), q0(Np,NeA), dqdt(Np,NeA), u(Np,NeA), v(Np,NeA), w(Np,NeA))
  allocate(idx(Np,NeA), qdev(Np,Ne))
  q = 0.0_RP; q0 = 0.0_RP; dqdt = 0.0_RP; u = 0.0_RP; v = 0.0_RP; w = 0.0_RP
  call dg_optr_kernel_setup(trim(cputype))

  ! the halo map, read back through update_halo
  nhalo_node = Nfp * 2 * (NeY*NeZ + NeX*NeZ + NeX*NeY)
  idx = 0.0_RP
  do ke = 1, Ne
    do n = 1, Np
      idx(n,ke) = real(n + (ke - 1) * Np, RP)
    end do
  end do
  call update_halo(idx)
  allocate(halo_src_map(nhalo_node))
  do ib = 1, nhalo_node
    halo_src_map(ib) = nint(idx(mod(ib - 1, Np) + 1, Ne + (ib - 1) / Np + 1))
  end do

  do ke = 1, Ne
    do n = 1, Np
      q(n,ke) = sin(2.0_RP*PI*pos_en(n,ke,1)) * sin(2.0_RP*PI*pos_en(n,ke,2)) &
              * sin(2.0_RP*PI*pos_en(n,ke,3))
      u(n,ke) = 1.0_RP
      v(n,ke) = 0.5_RP
      w(n,ke) = 0.25_RP
    end do
  end do
  call update_halo(q)
  call update_halo(u)
  call update_halo(v)
  call update_halo(w)

  ! the device solver on the reference's arrays
  call dgk_check(dgk_evaltype_id(dgk_cstr(devtype), evaltype), 'evaltype')
  call dgk_check(dgk_create(evaltype, int(Nq, c_int), int(Ne, c_int64_t), &
       int(NeA, c_int64_t), int(nhalo_node, c_int64_t), D1D, Lift_mat, VMapM, VMapP, &
       halo_src_map, normal_fn, Escale, Fscale, solver), 'dgk_create')
  nown = int(Np, c_int64_t) * Ne
  call dgk_check(dgk_set_field(solver, DGK_FIELD_Q, q, nown), 'set q')
  call dgk_check(dgk_set_field(solver, DGK_FIELD_U, u, nown), 'set u')
  call dgk_check(dgk_set_field(solver, DGK_FIELD_V, v, nown), 'set v')
  call dgk_check(dgk_set_field(solver, DGK_FIELD_W, w, nown), 'set w')

  do istep = 1, nstep
    call dgk_check(dgk_step(solver, real(dt, c_double), 1_c_int), 'dgk_step')
    call dgk_check(dgk_minmax(solver, dmin, dmax), 'dgk_minmax')
    do ke = 1, Ne
      q0(:,ke) = q(:,ke)
    end do
    do stage = 1, RK_nstage
      call update_halo(q)
      call advect3d_eq_cal_tend(dqdt, q, u, v, w, D1D, D1D_tr, Lift_mat, &
        VMapM, VMapP, normal_fn, Escale, Fscale, Nq, Np, NfpTot, Ne, NeA)
      do ke = 1, Ne
        q(:,ke) = rk_a(stage) * q0(:,ke) + rk_b(stage) * (q(:,ke) + dt * dqdt(:,ke))
      end do
    end do
    write(*,'(A,I8,4ES24.15)') 'STEP ', istep, dmin, dmax, minval(q(:,1:Ne)), maxval(q(:,1:Ne))
  end do
  call dgk_check(dgk_get_field(solver, DGK_FIELD_Q, qdev, nown), 'get q')
  call dgk_check(dgk_destroy(solver), 'dgk_destroy')
  write(*,'(A,ES24.15)') 'MAXREL ', maxval(abs(qdev - q(:,1:Ne))) / maxval(abs(q(:,1:Ne)))
  write(*,'(A,ES24.15)') 'RELL2 ', sqrt(sum((qdev - q(:,1:Ne))**2)) / sqrt(sum(q(:,1:Ne)**2))
end program ref_dgk_host
