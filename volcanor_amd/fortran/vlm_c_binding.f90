!! ISO_C_BINDING interfaces to libvlm (include/vlm_c.h).
!!
!! Every procedure of the C ABI is declared here with the reference's
!! argument meaning; the Fortran host (volcanor_host.f90) calls them from
!! type-bound procedures named after the reference's own
!! (classdef.f90, libCommon.f90).
module vlm_c
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: VLM_OK = 0
  integer(c_int), parameter :: VLM_FP64 = 0, VLM_MIXED = 1
  integer(c_int), parameter :: VLM_SET_WING = 0, VLM_SET_WING_BOUND = 1
  integer(c_int), parameter :: VLM_SET_WAKE = 2, VLM_SET_WAKE_PRED = 3
  integer(c_int), parameter :: VLM_WAKE_C = 0, VLM_WAKE_P = 1, VLM_WAKE_A = 2
  integer(c_int), parameter :: VLM_EDGE_LE = 0, VLM_EDGE_TE = 1
  integer(c_int), parameter :: VLM_TERM_NEGATE = 1
  integer(c_int), parameter :: VLM_VEL = 0, VLM_VEL_PREV = 1
  integer(c_int), parameter :: VLM_VEL_PRED = 2, VLM_VEL_STEP = 3

  type, bind(C) :: vlm_rotor_desc
    integer(c_int) :: nb, nc, ns
    integer(c_int) :: nNwake, nFwake
    integer(c_int) :: nbConvect
    integer(c_int) :: rollupStart, rollupEnd
    integer(c_int) :: rollupMin
    integer(c_int) :: suppressFwake
  end type vlm_rotor_desc

  interface
    integer(c_int) function vlm_create(ctx, device) bind(C, name='vlm_create')
      import :: c_int, c_ptr
      type(c_ptr), intent(out) :: ctx
      integer(c_int), value :: device
    end function

    integer(c_int) function vlm_destroy(ctx) bind(C, name='vlm_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    type(c_ptr) function vlm_last_error(ctx) bind(C, name='vlm_last_error')
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function vlm_set_precision(ctx, p) bind(C, name='vlm_set_precision')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: p
    end function

    integer(c_int) function vlm_set_bug_compat(ctx, b) bind(C, name='vlm_set_bug_compat')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: b
    end function

    integer(c_int) function vlm_synchronize(ctx) bind(C, name='vlm_synchronize')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function vlm_filament_vind(ctx, nfil, a, b, gam, rc, m, p, v, acc) &
        bind(C, name='vlm_filament_vind')
      import :: c_int, c_int64_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: nfil, m
      real(c_double), intent(in) :: a(3, *), b(3, *), gam(*), rc(*), p(3, *)
      real(c_double), intent(inout) :: v(3, *)
      integer(c_int), value :: acc
    end function

    integer(c_int) function vlm_rotor_create(ctx, desc, spanwiseCore, scv, ir) &
        bind(C, name='vlm_rotor_create')
      import :: c_int, c_ptr, c_double, vlm_rotor_desc
      type(c_ptr), value :: ctx
      type(vlm_rotor_desc), intent(in) :: desc
      real(c_double), value :: spanwiseCore
      real(c_double), intent(in) :: scv(*)
      integer(c_int), intent(out) :: ir
    end function

    integer(c_int) function vlm_rotor_set_wing(ctx, ir, corners, rc, gam) &
        bind(C, name='vlm_rotor_set_wing')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: ir
      real(c_double), intent(in) :: corners(*), rc(*), gam(*)
    end function

    integer(c_int) function vlm_rotor_set_axis(ctx, ir, shaft, hub) &
        bind(C, name='vlm_rotor_set_axis')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: ir
      real(c_double), intent(in) :: shaft(3), hub(3)
    end function

    integer(c_int) function vlm_rotor_set_init_wake_vel(ctx, ir, initWakeVel, initWakeVelNt) &
        bind(C, name='vlm_rotor_set_init_wake_vel')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: ir, initWakeVelNt
      real(c_double), value :: initWakeVel
    end function

    integer(c_int) function vlm_rotor_set_image(ctx, ir, source, imagePlane) &
        bind(C, name='vlm_rotor_set_image')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: ir, source, imagePlane
    end function

    integer(c_int) function vlm_rotor_mirror_gamma(ctx, ir) bind(C, name='vlm_rotor_mirror_gamma')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: ir
    end function

    integer(c_int) function vlm_rotor_mirror_wake(ctx, ir, wakeType) &
        bind(C, name='vlm_rotor_mirror_wake')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: ir, wakeType
    end function

    integer(c_int) function vlm_rotor_set_prescribed_wake(ctx, ir, prescWakeNt, prescWakeGenNt, &
        & deltaPsi) bind(C, name='vlm_rotor_set_prescribed_wake')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: ir, prescWakeNt, prescWakeGenNt
      real(c_double), value :: deltaPsi
    end function

    integer(c_int) function vlm_rotor_get_prescribed_wake(ctx, ir, wakeType, fc, gam, rc, built) &
        bind(C, name='vlm_rotor_get_prescribed_wake')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: ir, wakeType
      type(c_ptr), value :: fc, gam, rc
      integer(c_int), intent(out) :: built
    end function

    integer(c_int) function vlm_rotor_init_cursors(ctx, ir) bind(C, name='vlm_rotor_init_cursors')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: ir
    end function

    integer(c_int) function vlm_rotor_begin_step(ctx, ir, iter) bind(C, name='vlm_rotor_begin_step')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: ir, iter
    end function

    integer(c_int) function vlm_rotor_cursors(ctx, ir, rowNear, rowFar) &
        bind(C, name='vlm_rotor_cursors')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: ir
      integer(c_int), intent(out) :: rowNear, rowFar
    end function

    integer(c_int) function vlm_rotor_assignshed(ctx, ir, edge) bind(C, name='vlm_rotor_assignshed')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: ir, edge
    end function

    integer(c_int) function vlm_rotor_age_wake(ctx, ir, dt) bind(C, name='vlm_rotor_age_wake')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: ir
      real(c_double), value :: dt
    end function

    integer(c_int) function vlm_rotor_dissipate_wake(ctx, ir, dt, nu, avc, decay) &
        bind(C, name='vlm_rotor_dissipate_wake')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: ir
      real(c_double), value :: dt, nu, avc, decay
    end function

    integer(c_int) function vlm_rotor_strain_wake(ctx, ir) bind(C, name='vlm_rotor_strain_wake')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: ir
    end function

    integer(c_int) function vlm_rotor_burst_wake(ctx, ir, skewLimit, largeCore) &
        bind(C, name='vlm_rotor_burst_wake')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: ir
      real(c_double), value :: skewLimit, largeCore
    end function

    integer(c_int) function vlm_rotor_state_bytes(ctx, ir, bytes) &
        bind(C, name='vlm_rotor_state_bytes')
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: ir
      integer(c_int64_t), intent(out) :: bytes
    end function

    integer(c_int) function vlm_rotor_get_state(ctx, ir, buf, bytes) &
        bind(C, name='vlm_rotor_get_state')
      import :: c_int, c_int64_t, c_ptr, c_int8_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: ir
      integer(c_int8_t), intent(out) :: buf(*)
      integer(c_int64_t), value :: bytes
    end function

    integer(c_int) function vlm_rotor_set_state(ctx, ir, buf, bytes) &
        bind(C, name='vlm_rotor_set_state')
      import :: c_int, c_int64_t, c_ptr, c_int8_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: ir
      integer(c_int8_t), intent(in) :: buf(*)
      integer(c_int64_t), value :: bytes
    end function

    integer(c_int) function vlm_rotor_rollup(ctx, ir) bind(C, name='vlm_rotor_rollup')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: ir
    end function

    integer(c_int) function vlm_rotor_induced_velocity(ctx, ir, set, m, p, v, acc) &
        bind(C, name='vlm_rotor_induced_velocity')
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: ir, set
      integer(c_int64_t), value :: m
      real(c_double), intent(in) :: p(*)
      real(c_double), intent(inout) :: v(*)
      integer(c_int), value :: acc
    end function

    integer(c_int) function vlm_induced_velocity_terms(ctx, nterms, ir, set, flags, m, p, v, &
        & acc) bind(C, name='vlm_induced_velocity_terms')
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: nterms
      integer(c_int), intent(in) :: ir(*), set(*), flags(*)
      integer(c_int64_t), value :: m
      real(c_double), intent(in) :: p(*)
      real(c_double), intent(inout) :: v(*)
      integer(c_int), value :: acc
    end function

    integer(c_int) function vlm_convect(ctx, scheme, iter, dt) bind(C, name='vlm_convect')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: scheme, iter
      real(c_double), value :: dt
    end function

    integer(c_int) function vlm_rotor_calc_aic(ctx, ir, cp, ncap, A) bind(C, name='vlm_rotor_calc_aic')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: ir
      real(c_double), intent(in) :: cp(*), ncap(*)
      type(c_ptr), value :: A
    end function

    ! deferred variants: the outputs (passed as C addresses of TARGET
    ! arrays) are written by vlm_wait or by the next synchronising call
    integer(c_int) function vlm_rotor_solve_deferred(ctx, ir, rhs, gam, max_iter, tol, iters) &
        bind(C, name='vlm_rotor_solve_deferred')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: ir
      type(c_ptr), value :: rhs             ! C_NULL_PTR: the RHS of vlm_rotor_rhs
      type(c_ptr), value :: gam
      integer(c_int), value :: max_iter
      real(c_double), value :: tol
      type(c_ptr), value :: iters
    end function

    integer(c_int) function vlm_rotor_set_panels(ctx, ir, panel, section, blade) &
        bind(C, name='vlm_rotor_set_panels')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: ir
      real(c_double), intent(in) :: panel(*), section(*), blade(*)
    end function

    integer(c_int) function vlm_rotor_rhs(ctx, ir, nterms, irs, sets, flags, kin, rhs) &
        bind(C, name='vlm_rotor_rhs')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: ir, nterms
      integer(c_int), intent(in) :: irs(*), sets(*), flags(*)
      real(c_double), intent(in) :: kin(*)
      type(c_ptr), value :: rhs
    end function

    integer(c_int) function vlm_rotor_mirror_velcp(ctx, ir) bind(C, name='vlm_rotor_mirror_velcp')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: ir
    end function

    integer(c_int) function vlm_rotor_loads(ctx, ir, nterms, irs, sets, flags, density, dt, &
        & omegaSign, spanwiseLiftSwitch, out) bind(C, name='vlm_rotor_loads')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: ir, nterms
      integer(c_int), intent(in) :: irs(*), sets(*), flags(*)
      real(c_double), value :: density, dt, omegaSign
      integer(c_int), value :: spanwiseLiftSwitch
      type(c_ptr), value :: out
    end function

    integer(c_int) function vlm_induced_velocity_terms_deferred(ctx, nterms, ir, set, flags, &
        & m, p, v, acc) bind(C, name='vlm_induced_velocity_terms_deferred')
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: nterms
      integer(c_int), intent(in) :: ir(*), set(*), flags(*)
      integer(c_int64_t), value :: m
      real(c_double), intent(in) :: p(*)
      type(c_ptr), value :: v
      integer(c_int), value :: acc
    end function

    integer(c_int) function vlm_wait(ctx) bind(C, name='vlm_wait')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function vlm_rotor_solve(ctx, ir, rhs, gam, max_iter, tol, iters) &
        bind(C, name='vlm_rotor_solve')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: ir
      real(c_double), intent(in) :: rhs(*)
      real(c_double), intent(out) :: gam(*)
      integer(c_int), value :: max_iter
      real(c_double), value :: tol
      integer(c_int), intent(out) :: iters
    end function

    integer(c_int) function vlm_rotor_get_wake(ctx, ir, wakeType, nodes, le, gam, rc4, &
        far_fc, far_gam, far_rc) bind(C, name='vlm_rotor_get_wake')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: ir, wakeType
      type(c_ptr), value :: nodes, le, gam, rc4, far_fc, far_gam, far_rc
    end function

    integer(c_int) function vlm_set_predictor(ctx, mode, tol, stride, warmup, iters, lref) &
        bind(C, name='vlm_set_predictor')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: mode
      real(c_double), value :: tol
      integer(c_int), value :: stride, warmup, iters
      real(c_double), value :: lref
    end function

    integer(c_int) function vlm_predictor_stats(ctx, full, learned, fallbacks, last_err, &
        last_loss) bind(C, name='vlm_predictor_stats')
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), intent(out) :: full, learned, fallbacks
      real(c_double), intent(out) :: last_err, last_loss
    end function

    integer(c_int) function vlm_predictor_errors(ctx, mlp_rel_rms, base_rel_rms, judged) &
        bind(C, name='vlm_predictor_errors')
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: mlp_rel_rms, base_rel_rms
      integer(c_int64_t), intent(out) :: judged
    end function

    integer(c_int) function vlm_profile_get_lu(ctx, ms, factorisations, flops) &
        bind(C, name='vlm_profile_get_lu')
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: ms, flops
      integer(c_int64_t), intent(out) :: factorisations
    end function

    integer(c_int) function vlm_profile_enable(ctx, on) bind(C, name='vlm_profile_enable')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: on
    end function

    integer(c_int) function vlm_profile_reset(ctx) bind(C, name='vlm_profile_reset')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function vlm_profile_get(ctx, ms, launches, pairs) bind(C, name='vlm_profile_get')
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: ms, pairs
      integer(c_int64_t), intent(out) :: launches
    end function

    integer(c_int) function vlm_profile_get_all(ctx, pairs, launches) &
        bind(C, name='vlm_profile_get_all')
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: pairs
      integer(c_int64_t), intent(out) :: launches
    end function
  end interface

contains

  subroutine vlm_check(ctx, rc, where)
    !! Turns a non-zero status into a stop of the run with the library
    !! message (host_fail).
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    character(len=512) :: text
    character(len=700) :: line
    integer :: i

    if (rc == VLM_OK) return
    text = ''
    if (c_associated(ctx)) then
      call c_f_pointer(vlm_last_error(ctx), msg, [512])
      do i = 1, 512
        if (msg(i) == c_null_char) exit
        text(i:i) = msg(i)
      enddo
    endif
    write (line, '(A,A,A,I0,A,A)') 'ERROR in ', where, ' (code ', rc, '): ', trim(text)
    call host_fail(trim(line))
  end subroutine vlm_check

  subroutine host_fail(msg)
    !! Stops the run with msg: the caller of volcanor_run gets status 1 and
    !! the message (volcanor_last_error); a program stops with status 1, as
    !! the reference's `error stop` does (host_guard.c).
    character(len=*), intent(in) :: msg
    interface
      subroutine volcanor_host_fail(m, n) bind(C, name='volcanor_host_fail')
        import :: c_char, c_int
        character(kind=c_char), intent(in) :: m(*)
        integer(c_int), value :: n
      end subroutine
    end interface
    call volcanor_host_fail(msg, int(len(msg), c_int))
  end subroutine host_fail

end module vlm_c
