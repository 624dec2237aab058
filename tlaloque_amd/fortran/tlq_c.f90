! tlq_c.f90 -- ISO_C_BINDING interface to libtlq (include/tlq.h).
!
! The reference's hot path is a set of implicit-interface Fortran subroutines
! working on the module arrays of source/globals.f90:5-11 (calc_prims,
! time_step, step: source/HydroCore.f90:1,113,170; HLLCFluxes, Godunov:
! source/HLLC.f90:64,1; borders: source/boundaries.f90:1; cooling:
! source/cooling.f90:9).  On the GPU those arrays live in device memory, so a
! Fortran driver holds a context handle and names the arrays by TLQ_BUF_* ids;
! host arrays cross the boundary only at upload/download, in the reference's
! layout real(8) U(5, -1:nx+2, -1:ny+2, -1:nz+2).
!
! Every function returns TLQ_OK (0) or a negative TLQ_ERR_* code;
! tlq_require stops the program with the library's message, the way the
! reference stops on a bad limiter (HLLC.f90:630-632).
module tlq_c
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: TLQ_OK = 0
  integer(c_int), parameter :: TLQ_ERR_ARG = -1, TLQ_ERR_HIP = -2, TLQ_ERR_NOMEM = -3
  integer(c_int), parameter :: TLQ_ERR_UNSUPPORTED = -4, TLQ_ERR_STATE = -5, TLQ_ERR_NONFINITE = -6

  ! source/constants.f90:20-31
  integer(c_int), parameter :: TLQ_BC_OUTFLOW = 1, TLQ_BC_REFLECTIVE = 2, TLQ_BC_PERIODIC = 3
  integer(c_int), parameter :: TLQ_BC_INFLOW = 4
  integer(c_int), parameter :: TLQ_LIM_NONE = 1, TLQ_LIM_VANLEER = 2, TLQ_LIM_MINMOD = 3
  integer(c_int), parameter :: TLQ_LIM_UMIST = 4, TLQ_LIM_WOODWARD = 5, TLQ_LIM_SUPERBEE = 6

  integer(c_int), parameter :: TLQ_SOLVER_HLLC = 0, TLQ_SOLVER_HLL = 1
  integer(c_int), parameter :: TLQ_FP64 = 0, TLQ_FP32 = 1
  integer(c_int), parameter :: TLQ_COOL_OFF = 0, TLQ_COOL_EXACT = 1, TLQ_COOL_SURROGATE = 2
  integer(c_int), parameter :: TLQ_COOL_SUBCYCLED = 3, TLQ_COOL_INTEGRATED = 4
  integer(c_int), parameter :: TLQ_SG_OFF = 0, TLQ_SG_FIXED = 1, TLQ_SG_MULTIPOLE = 2

  integer(c_int), parameter :: TLQ_BUF_U = 0, TLQ_BUF_UP = 1, TLQ_BUF_UPP = 2, TLQ_BUF_UPRIM = 3
  integer(c_int), parameter :: TLQ_BUF_F = 4, TLQ_BUF_G = 5, TLQ_BUF_H = 6, TLQ_BUF_S = 7

  integer(c_int), parameter :: TLQ_MLP_NPARAMS = 4481

  ! struct tlq_params: user/parameters.f90:10-49 as run-time values
  type, bind(C) :: tlq_params_t
     integer(c_int) :: nx, ny, nz
     real(c_double) :: xl, xr, yl, yr, zl, zr
     real(c_double) :: gamma, mu0, cfl, eta, cooling_limit
     integer(c_int) :: limiter
     integer(c_int) :: bc(6)          ! LB, RB, DB, TB, FB, BB
     integer(c_int) :: solver, precision, cooling
     real(c_double) :: softening
     integer(c_int) :: self_gravity   ! TLQ_SG_*
     integer(c_int) :: mg_levels      ! lvlm; 0: as many as the mesh halves
     real(c_double) :: gconst         ! Gconst
     real(c_double) :: mg_tol         ! V-cycles while Dif > mg_tol; 0: the reference's 10
  end type tlq_params_t

  ! struct tlq_wind: spherical_wind_type, source/winds.f90:5-13
  type, bind(C) :: tlq_wind_t
     real(c_double) :: xc, yc, zc, vx, vy, vz, radius, mdot, vinf, temp, mu
  end type tlq_wind_t

  interface
     integer(c_int) function tlq_create(params, ctx) bind(C, name="tlq_create")
       import :: c_int, c_ptr, tlq_params_t
       type(tlq_params_t), intent(in) :: params
       type(c_ptr), intent(out) :: ctx
     end function tlq_create

     integer(c_int) function tlq_destroy(ctx) bind(C, name="tlq_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function tlq_destroy

     type(c_ptr) function tlq_last_error() bind(C, name="tlq_last_error")
       import :: c_ptr
     end function tlq_last_error

     integer(c_int) function tlq_set_winds(ctx, winds, n) bind(C, name="tlq_set_winds")
       import :: c_int, c_ptr, tlq_wind_t
       type(c_ptr), value :: ctx
       type(tlq_wind_t), intent(in) :: winds(*)
       integer(c_int), value :: n
     end function tlq_set_winds

     integer(c_int) function tlq_set_point_masses(ctx, xyzm, n) bind(C, name="tlq_set_point_masses")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: xyzm(*)
       integer(c_int), value :: n
     end function tlq_set_point_masses

     integer(c_int) function tlq_set_cooling_table(ctx, temp, log_lambda, n) &
          bind(C, name="tlq_set_cooling_table")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: temp(*), log_lambda(*)
       integer(c_int), value :: n
     end function tlq_set_cooling_table

     integer(c_int) function tlq_upload(ctx, buf, host) bind(C, name="tlq_upload")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: buf
       real(c_double), intent(in) :: host(*)
     end function tlq_upload

     integer(c_int) function tlq_download(ctx, buf, host) bind(C, name="tlq_download")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: buf
       real(c_double), intent(out) :: host(*)
     end function tlq_download

     integer(c_int) function tlq_upload_phi(ctx, phi) bind(C, name="tlq_upload_phi")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: phi(*)
     end function tlq_upload_phi

     integer(c_int) function tlq_download_phi(ctx, phi) bind(C, name="tlq_download_phi")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: phi(*)
     end function tlq_download_phi

     integer(c_int) function tlq_gravity_solve(ctx, cycles, dif) bind(C, name="tlq_gravity_solve")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), intent(out) :: cycles
       real(c_double), intent(out) :: dif
     end function tlq_gravity_solve

     integer(c_int) function tlq_gravity_info(ctx, levels, cycles, dif, it) &
          bind(C, name="tlq_gravity_info")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), intent(out) :: levels, cycles, it
       real(c_double), intent(out) :: dif
     end function tlq_gravity_info

     integer(c_int) function tlq_calc_prims(ctx, src) bind(C, name="tlq_calc_prims")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: src
     end function tlq_calc_prims

     integer(c_int) function tlq_time_step(ctx, dt) bind(C, name="tlq_time_step")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: dt
     end function tlq_time_step

     integer(c_int) function tlq_fluxes(ctx, src, choice) bind(C, name="tlq_fluxes")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: src, choice
     end function tlq_fluxes

     integer(c_int) function tlq_godunov(ctx, dst, dt) bind(C, name="tlq_godunov")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: dst
       real(c_double), value :: dt
     end function tlq_godunov

     integer(c_int) function tlq_borders(ctx, buf, order) bind(C, name="tlq_borders")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: buf, order
     end function tlq_borders

     integer(c_int) function tlq_step(ctx, t, it, dt) bind(C, name="tlq_step")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(inout) :: t
       integer(c_int), intent(inout) :: it
       real(c_double), value :: dt
     end function tlq_step

     integer(c_int) function tlq_cooling(ctx, dt) bind(C, name="tlq_cooling")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: dt
     end function tlq_cooling

     integer(c_int) function tlq_init_state(ctx) bind(C, name="tlq_init_state")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function tlq_init_state

     integer(c_int) function tlq_advance(ctx, dt, t, it) bind(C, name="tlq_advance")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: dt, t
       integer(c_int), intent(out) :: it
     end function tlq_advance

     integer(c_int) function tlq_advance_n(ctx, n) bind(C, name="tlq_advance_n")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: n
     end function tlq_advance_n

     integer(c_int) function tlq_get_time(ctx, dt, t, it) bind(C, name="tlq_get_time")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: dt, t
       integer(c_int), intent(out) :: it
     end function tlq_get_time

     integer(c_int) function tlq_resume(ctx, t, it, fill_ghosts) bind(C, name="tlq_resume")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: t
       integer(c_int), value :: it, fill_ghosts
     end function tlq_resume

     integer(c_int) function tlq_sync(ctx) bind(C, name="tlq_sync")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function tlq_sync

     integer(c_int) function tlq_check_state(ctx, bad) bind(C, name="tlq_check_state")
       import :: c_int, c_ptr, c_long_long
       type(c_ptr), value :: ctx
       integer(c_long_long), intent(out) :: bad
     end function tlq_check_state

     integer(c_int) function tlq_fallback_fraction(ctx, frac) bind(C, name="tlq_fallback_fraction")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: frac
     end function tlq_fallback_fraction

     integer(c_int) function tlq_cooling_counts(ctx, queued, fallbacks) bind(C, name="tlq_cooling_counts")
       import :: c_int, c_ptr, c_long_long
       type(c_ptr), value :: ctx
       integer(c_long_long), intent(out) :: queued, fallbacks
     end function tlq_cooling_counts

     integer(c_int) function tlq_surrogate_init(ctx, seed) bind(C, name="tlq_surrogate_init")
       import :: c_int, c_ptr, c_long_long
       type(c_ptr), value :: ctx
       integer(c_long_long), value :: seed
     end function tlq_surrogate_init

     integer(c_int) function tlq_surrogate_set_params(ctx, theta, n) &
          bind(C, name="tlq_surrogate_set_params")
       import :: c_int, c_ptr, c_float
       type(c_ptr), value :: ctx
       real(c_float), intent(in) :: theta(*)
       integer(c_int), value :: n
     end function tlq_surrogate_set_params

     integer(c_int) function tlq_surrogate_get_params(ctx, theta, n) &
          bind(C, name="tlq_surrogate_get_params")
       import :: c_int, c_ptr, c_float
       type(c_ptr), value :: ctx
       real(c_float), intent(out) :: theta(*)
       integer(c_int), value :: n
     end function tlq_surrogate_get_params

     integer(c_int) function tlq_surrogate_generate(ctx, nsamples, seed) &
          bind(C, name="tlq_surrogate_generate")
       import :: c_int, c_ptr, c_long_long
       type(c_ptr), value :: ctx
       integer(c_long_long), value :: nsamples, seed
     end function tlq_surrogate_generate

     integer(c_int) function tlq_surrogate_samples(ctx, first, n, x, y) &
          bind(C, name="tlq_surrogate_samples")
       import :: c_int, c_ptr, c_long_long, c_float
       type(c_ptr), value :: ctx
       integer(c_long_long), value :: first, n
       real(c_float), intent(out) :: x(*), y(*)
     end function tlq_surrogate_samples

     integer(c_int) function tlq_surrogate_train(ctx, steps, batch, lr_max, lr_min, final_loss) &
          bind(C, name="tlq_surrogate_train")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: steps, batch
       real(c_double), value :: lr_max, lr_min
       real(c_double), intent(out) :: final_loss
     end function tlq_surrogate_train

     integer(c_int) function tlq_surrogate_eval(ctx, x, n, out) bind(C, name="tlq_surrogate_eval")
       import :: c_int, c_ptr, c_long_long, c_float
       type(c_ptr), value :: ctx
       real(c_float), intent(in) :: x(*)
       integer(c_long_long), value :: n
       real(c_float), intent(out) :: out(*)
     end function tlq_surrogate_eval

     integer(c_int) function tlq_surrogate_grad(ctx, x, y, n, grad, loss) &
          bind(C, name="tlq_surrogate_grad")
       import :: c_int, c_ptr, c_float, c_double
       type(c_ptr), value :: ctx
       real(c_float), intent(in) :: x(*), y(*)
       integer(c_int), value :: n
       real(c_double), intent(out) :: grad(*), loss
     end function tlq_surrogate_grad

     integer(c_int) function tlq_timer_start(ctx) bind(C, name="tlq_timer_start")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function tlq_timer_start

     integer(c_int) function tlq_timer_stop(ctx, ms) bind(C, name="tlq_timer_stop")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: ms
     end function tlq_timer_stop

     integer(c_int) function tlq_kernel_timing(ctx, enable) bind(C, name="tlq_kernel_timing")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: enable
     end function tlq_kernel_timing

     integer(c_int) function tlq_kernel_stats(ctx, idx, name, name_len, total_ms, launches) &
          bind(C, name="tlq_kernel_stats")
       import :: c_int, c_ptr, c_char, c_double, c_long_long
       type(c_ptr), value :: ctx
       integer(c_int), value :: idx
       character(kind=c_char), intent(out) :: name(*)
       integer(c_int), value :: name_len
       real(c_double), intent(out) :: total_ms
       integer(c_long_long), intent(out) :: launches
     end function tlq_kernel_stats

     integer(c_int) function tlq_copy_probe(ctx, bytes, iters, ms) bind(C, name="tlq_copy_probe")
       import :: c_int, c_ptr, c_long_long, c_double
       type(c_ptr), value :: ctx
       integer(c_long_long), value :: bytes
       integer(c_int), value :: iters
       real(c_double), intent(out) :: ms
     end function tlq_copy_probe

     integer(c_int) function tlq_layout(ctx, cells, pitch_x, origin_x) bind(C, name="tlq_layout")
       import :: c_int, c_ptr, c_long_long
       type(c_ptr), value :: ctx
       integer(c_long_long), intent(out) :: cells
       integer(c_int), intent(out) :: pitch_x, origin_x
     end function tlq_layout

     integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
       import :: c_size_t, c_ptr
       type(c_ptr), value :: s
     end function c_strlen
  end interface

contains

  ! The library's last error message as a Fortran string.
  function tlq_error_message() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: n, q
    p = tlq_last_error()
    if (.not. c_associated(p)) then
       msg = ''
       return
    end if
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate(character(len=n) :: msg)
    do q = 1, n
       msg(q:q) = chars(q)
    end do
  end function tlq_error_message

  ! Stop with the library's message when rc is not TLQ_OK.
  subroutine tlq_require(rc, what)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: what
    if (rc /= TLQ_OK) then
       write(*, '(a,a,a,i0,a,a)') 'tlq: ', what, ' failed (', rc, '): ', tlq_error_message()
       error stop 1
    end if
  end subroutine tlq_require

end module tlq_c
