!> Reference-side harness (TEST INFRASTRUCTURE ONLY).
!!
!! Links the reference modules (mod_common, mod_mesh, mod_dg_optr_kernel[_opt1],
!! mod_advect3d_eq), compiled by oracle/Makefile straight from the reference
!! sources, and drives them on caller-supplied inputs.  The time loop below
!! restates the reference driver (main.f90:51-80) with the initial condition
!! and velocity read from a file instead of the built-in sinusoid
!! (main.f90:146-157), so that the GPU path, the C oracle and the reference see
!! bit-identical inputs.
!!
!! Usage (cwd must contain operator_data/, as for the reference binary):
!!   ref_harness MODE NeX NeY NeZ p OPTTYPE nstep dt infile outfile
!!   MODE = run  : nstep SSP-RK3 steps; outfile <- q(Np,Ne), then min/max per step
!!   MODE = tend : update_halo + one advect3d_eq_cal_tend; outfile <- dqdt(Np,Ne)
!!   MODE = time : like run, writes nothing; prints the loop time
!! infile holds q, u, v, w as four (Np,Ne) float64 arrays (stream access).
!! stdout: "LOOP_SECONDS <t>" (wall time of the loop, main.f90:164->174 span).
program ref_harness
  use mod_common, only: RP, &
    RK_nstage => RK3s3oSSP_nstage, rk_a => RK3s3oSSP_rk_a, rk_b => RK3s3oSSP_rk_b
  use mod_mesh, only: mesh_setup, update_halo, &
    Nq, Np, NfpTot, Ne, NeA, D1D, D1D_tr, Lift_mat, VMapM, VMapP, &
    normal_fn, Escale, Fscale
  use mod_advect3d_eq, only: advect3d_eq_cal_tend
  use mod_dg_optr_kernel, only: dg_optr_kernel_setup
  implicit none

  character(len=256) :: arg, mode, opttype, infile, outfile
  integer :: NeX, NeY, NeZ, p, nstep, istep, stage, fid, ke
  real(RP) :: dt
  real(RP), allocatable :: q(:,:), q0(:,:), dqdt(:,:), u(:,:), v(:,:), w(:,:)
  real(RP), allocatable :: tmp(:,:), trace(:,:)
  integer(8) :: c0, c1, crate

  if (command_argument_count() < 10) then
    write(*,'(A)') 'usage: ref_harness MODE NeX NeY NeZ p OPTTYPE nstep dt infile outfile'
    error stop 1
  end if
  call get_command_argument(1, mode)
  call get_command_argument(2, arg); read(arg,*) NeX
  call get_command_argument(3, arg); read(arg,*) NeY
  call get_command_argument(4, arg); read(arg,*) NeZ
  call get_command_argument(5, arg); read(arg,*) p
  call get_command_argument(6, opttype)
  call get_command_argument(7, arg); read(arg,*) nstep
  call get_command_argument(8, arg); read(arg,*) dt
  call get_command_argument(9, infile)
  call get_command_argument(10, outfile)

  call mesh_setup(NeX, NeY, NeZ, p, 1.0_RP, 1.0_RP, 1.0_RP)
  allocate(q(Np,NeA), q0(Np,NeA), dqdt(Np,NeA), u(Np,NeA), v(Np,NeA), w(Np,NeA))
  allocate(tmp(Np,Ne), trace(2,max(nstep,1)))
  q = 0.0_RP; q0 = 0.0_RP; dqdt = 0.0_RP; u = 0.0_RP; v = 0.0_RP; w = 0.0_RP
  call dg_optr_kernel_setup(trim(opttype))

  open(newunit=fid, file=trim(infile), access='stream', form='unformatted', &
       status='old', action='read')
  read(fid) tmp; q(:,1:Ne) = tmp
  read(fid) tmp; u(:,1:Ne) = tmp
  read(fid) tmp; v(:,1:Ne) = tmp
  read(fid) tmp; w(:,1:Ne) = tmp
  close(fid)

  call update_halo(q)
  call update_halo(u)
  call update_halo(v)
  call update_halo(w)

  call system_clock(c0, crate)
  if (trim(mode) == 'tend') then
    call advect3d_eq_cal_tend(dqdt, q, u, v, w, D1D, D1D_tr, Lift_mat, &
      VMapM, VMapP, normal_fn, Escale, Fscale, Nq, Np, NfpTot, Ne, NeA)
  else
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
      if (trim(mode) == 'run') then
        trace(1,istep) = minval(q(:,1:Ne))
        trace(2,istep) = maxval(q(:,1:Ne))
      end if
    end do
  end if
  call system_clock(c1)
  write(*,'(A,ES24.15)') 'LOOP_SECONDS ', real(c1 - c0, RP) / real(crate, RP)

  if (trim(mode) == 'tend') then
    open(newunit=fid, file=trim(outfile), access='stream', form='unformatted', &
         status='replace', action='write')
    tmp = dqdt(:,1:Ne)
    write(fid) tmp
    close(fid)
  else if (trim(mode) == 'run') then
    open(newunit=fid, file=trim(outfile), access='stream', form='unformatted', &
         status='replace', action='write')
    tmp = q(:,1:Ne)
    write(fid) tmp
    write(fid) trace(:,1:nstep)
    close(fid)
  end if
end program ref_harness
