!> Fortran interfaces of libdgk (include/dgk.h) via iso_c_binding.
!!
!! The C ABI takes the reference's own arrays as they are declared in the
!! Fortran host (column-major real64, default integer, 1-based indices), so a
!! Fortran caller passes them straight through:
!!   * arrays as assumed-size dummies, passed by reference;
!!   * scalars by value;
!!   * strings as NUL-terminated character(kind=c_char) arrays.
!! Every entry point returns a status (0 = success); dgk_check stops the
!! program with the library's message otherwise, the way the reference host
!! stops on a bad configuration (main.f90:117-122).
module mod_dgk
  use, intrinsic :: iso_c_binding
  implicit none
  private

  integer(c_int), public, parameter :: DGK_OK = 0
  integer(c_int), public, parameter :: DGK_EVAL_GENERAL = 1, DGK_EVAL_OPT1 = 2
  integer(c_int), public, parameter :: DGK_EVAL_HIP = 3, DGK_EVAL_HIP_MP = 4
  integer(c_int), public, parameter :: DGK_FIELD_Q = 0, DGK_FIELD_U = 1
  integer(c_int), public, parameter :: DGK_FIELD_V = 2, DGK_FIELD_W = 3
  integer(c_int), public, parameter :: DGK_FIELD_DQDT = 4

  public :: dgk_check, dgk_cstr
  public :: dgk_last_error, dgk_evaltype_id, dgk_element_sizes, dgk_load_operators
  public :: dgk_mesh_sizes, dgk_mesh_build, dgk_create, dgk_create_structured
  public :: dgk_destroy, dgk_set_field, dgk_get_field, dgk_update_halo, dgk_copy_q0
  public :: dgk_cal_elembnd_flux, dgk_cal_dqdt, dgk_cal_tend, dgk_rk_update, dgk_step
  public :: dgk_minmax, dgk_synchronize, dgk_set_kernel_timing, dgk_timer_report
  public :: dgk_device_count

  interface
    function dgk_last_error() bind(C, name="dgk_last_error") result(msg)
      import :: c_ptr
      type(c_ptr) :: msg
    end function
    function dgk_device_count(count) bind(C, name="dgk_device_count") result(st)
      import :: c_int
      integer(c_int) :: count
      integer(c_int) :: st
    end function
    function dgk_evaltype_id(name, id) bind(C, name="dgk_evaltype_id") result(st)
      import :: c_int, c_char
      character(kind=c_char) :: name(*)
      integer(c_int) :: id
      integer(c_int) :: st
    end function
    function dgk_element_sizes(p, Nq, Np, Nfp, NfpTot) bind(C, name="dgk_element_sizes") result(st)
      import :: c_int
      integer(c_int), value :: p
      integer(c_int) :: Nq, Np, Nfp, NfpTot
      integer(c_int) :: st
    end function
    function dgk_load_operators(dir, p, x1D, D1D, Lift_mat) bind(C, name="dgk_load_operators") result(st)
      import :: c_int, c_char, c_double
      character(kind=c_char) :: dir(*)
      integer(c_int), value :: p
      real(c_double) :: x1D(*), D1D(*), Lift_mat(*)
      integer(c_int) :: st
    end function
    function dgk_mesh_sizes(NeX, NeY, NeZ, p, Ne, NeA, NhaloNode) bind(C, name="dgk_mesh_sizes") result(st)
      import :: c_int, c_int64_t
      integer(c_int), value :: NeX, NeY, NeZ, p
      integer(c_int64_t) :: Ne, NeA, NhaloNode
      integer(c_int) :: st
    end function
    function dgk_mesh_build(NeX, NeY, NeZ, p, x1D, dom_x, dom_y, dom_z, VMapM, VMapP, &
        halo_src_map, normal_fn, Escale, Fscale, pos_en) bind(C, name="dgk_mesh_build") result(st)
      import :: c_int, c_double
      integer(c_int), value :: NeX, NeY, NeZ, p
      real(c_double) :: x1D(*)
      real(c_double), value :: dom_x, dom_y, dom_z
      integer(c_int) :: VMapM(*), VMapP(*), halo_src_map(*)
      real(c_double) :: normal_fn(*), Escale(*), Fscale(*), pos_en(*)
      integer(c_int) :: st
    end function
    function dgk_create(evaltype, Nq, Ne, NeA, NhaloNode, D1D, Lift_mat, VMapM, VMapP, &
        halo_src_map, normal_fn, Escale, Fscale, solver) bind(C, name="dgk_create") result(st)
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: evaltype, Nq
      integer(c_int64_t), value :: Ne, NeA, NhaloNode
      real(c_double) :: D1D(*), Lift_mat(*)
      integer(c_int) :: VMapM(*), VMapP(*), halo_src_map(*)
      real(c_double) :: normal_fn(*), Escale(*), Fscale(*)
      type(c_ptr) :: solver
      integer(c_int) :: st
    end function
    function dgk_create_structured(evaltype, NeX, NeY, NeZ, p, dir, dom_x, dom_y, dom_z, &
        solver) bind(C, name="dgk_create_structured") result(st)
      import :: c_int, c_char, c_double, c_ptr
      integer(c_int), value :: evaltype, NeX, NeY, NeZ, p
      character(kind=c_char) :: dir(*)
      real(c_double), value :: dom_x, dom_y, dom_z
      type(c_ptr) :: solver
      integer(c_int) :: st
    end function
    function dgk_destroy(solver) bind(C, name="dgk_destroy") result(st)
      import :: c_int, c_ptr
      type(c_ptr), value :: solver
      integer(c_int) :: st
    end function
    function dgk_set_field(solver, which, f, n) bind(C, name="dgk_set_field") result(st)
      import :: c_int, c_int64_t, c_double, c_ptr
      type(c_ptr), value :: solver
      integer(c_int), value :: which
      real(c_double) :: f(*)
      integer(c_int64_t), value :: n
      integer(c_int) :: st
    end function
    function dgk_get_field(solver, which, f, n) bind(C, name="dgk_get_field") result(st)
      import :: c_int, c_int64_t, c_double, c_ptr
      type(c_ptr), value :: solver
      integer(c_int), value :: which
      real(c_double) :: f(*)
      integer(c_int64_t), value :: n
      integer(c_int) :: st
    end function
    function dgk_update_halo(solver, which) bind(C, name="dgk_update_halo") result(st)
      import :: c_int, c_ptr
      type(c_ptr), value :: solver
      integer(c_int), value :: which
      integer(c_int) :: st
    end function
    function dgk_copy_q0(solver) bind(C, name="dgk_copy_q0") result(st)
      import :: c_int, c_ptr
      type(c_ptr), value :: solver
      integer(c_int) :: st
    end function
    function dgk_cal_elembnd_flux(solver, flux) bind(C, name="dgk_cal_elembnd_flux") result(st)
      import :: c_int, c_ptr
      type(c_ptr), value :: solver
      type(c_ptr), value :: flux
      integer(c_int) :: st
    end function
    function dgk_cal_dqdt(solver, flux) bind(C, name="dgk_cal_dqdt") result(st)
      import :: c_int, c_ptr
      type(c_ptr), value :: solver
      type(c_ptr), value :: flux
      integer(c_int) :: st
    end function
    function dgk_cal_tend(solver, dqdt) bind(C, name="dgk_cal_tend") result(st)
      import :: c_int, c_ptr
      type(c_ptr), value :: solver
      type(c_ptr), value :: dqdt
      integer(c_int) :: st
    end function
    function dgk_rk_update(solver, stage, dt) bind(C, name="dgk_rk_update") result(st)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: solver
      integer(c_int), value :: stage
      real(c_double), value :: dt
      integer(c_int) :: st
    end function
    function dgk_step(solver, dt, nstep) bind(C, name="dgk_step") result(st)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: solver
      real(c_double), value :: dt
      integer(c_int), value :: nstep
      integer(c_int) :: st
    end function
    function dgk_minmax(solver, qmin, qmax) bind(C, name="dgk_minmax") result(st)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: solver
      real(c_double) :: qmin, qmax
      integer(c_int) :: st
    end function
    function dgk_synchronize(solver) bind(C, name="dgk_synchronize") result(st)
      import :: c_int, c_ptr
      type(c_ptr), value :: solver
      integer(c_int) :: st
    end function
    function dgk_set_kernel_timing(solver, on) bind(C, name="dgk_set_kernel_timing") result(st)
      import :: c_int, c_ptr
      type(c_ptr), value :: solver
      integer(c_int), value :: on
      integer(c_int) :: st
    end function
    function dgk_timer_report(solver, seconds, counts) bind(C, name="dgk_timer_report") result(st)
      import :: c_int, c_int64_t, c_double, c_ptr
      type(c_ptr), value :: solver
      real(c_double) :: seconds(*)
      integer(c_int64_t) :: counts(*)
      integer(c_int) :: st
    end function
  end interface

contains

  !> NUL-terminated copy of a Fortran string (trailing blanks removed).
  function dgk_cstr(s) result(c)
    character(len=*), intent(in) :: s
    character(kind=c_char) :: c(len_trim(s) + 1)
    integer :: i
    do i = 1, len_trim(s)
      c(i) = s(i:i)
    end do
    c(len_trim(s) + 1) = c_null_char
  end function dgk_cstr

  !> Stop with the library's message when a call failed.
  subroutine dgk_check(st, what)
    use, intrinsic :: iso_fortran_env, only: error_unit
    integer(c_int), intent(in) :: st
    character(len=*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    character(len=512) :: text
    integer :: i
    if (st == DGK_OK) return
    text = ''
    call c_f_pointer(dgk_last_error(), msg, [512])
    do i = 1, 512
      if (msg(i) == c_null_char) exit
      text(i:i) = msg(i)
    end do
    write(error_unit, '(A,A,I0,A,A)') trim(what), ': status ', st, ': ', trim(text)
    flush(error_unit)
    error stop 2
  end subroutine dgk_check

end module mod_dgk
