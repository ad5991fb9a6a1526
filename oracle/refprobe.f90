!! refprobe: test infrastructure, not product code.
!!
!! Built by `make -C oracle ref` against the reference's own objects
!! (src/libMath_noBLAS.f90, classdef.f90, libCommon.f90, ... compiled where
!! they lie).  It runs the reference's free-wake time step on the case in
!! the current directory and writes the wing and wake state after the
!! initial solution and after every step, in full precision, to
!! Results/refprobe.txt.  The oracle's wake functions (rotor_assignshed,
!! age, dissipate, strain, rollup/shift, convectwake and the AM2 scheme
!! algebra) are then pinned against these states node by node, with the
!! reference's own wing circulation as input.
!!
!! The step follows src/main.f90 (initial solution 117-234, time loop
!! 400-1452) for the options it accepts: lifting surfaces without images,
!! fdScheme 3 (AM2), wake dissipation and strain switches, no slow start,
!! no sub-iterations, no flap dynamics, no custom trajectory, no loads.
program refprobe
  use libMath, only: matmulAX, cross_product, eps
  use libCommon
  implicit none
  integer, parameter :: dp = kind(1.d0)
  integer :: u, k

  call readConfig('config.nml', 'Results/config.nml')
  if (switches%fdScheme /= 3) error stop 'refprobe: fdScheme 3 (AM2) only'
  if (switches%slowStart /= 0 .or. switches%ntSub /= 0 .or. switches%ntSubInit /= 0) &
    error stop 'refprobe: no slow start, no sub-iterations'
  allocate (rotor(nr))
  do ir = 1, nr
    write (rotorChar, '(I0.2)') ir
    call rotor(ir)%readGeom('geom'//rotorChar//'.nml', 'Results/geom'//rotorChar//'.nml')
  enddo
  do ir = 1, nr
    if (rotor(ir)%surfaceType < 0) error stop 'refprobe: image rotors not supported'
    call rotor(ir)%init(ir, density, dt, nt, switches)
    if (rotor(ir)%bladeDynamicsSwitch /= 0) error stop 'refprobe: no blade dynamics'
  enddo
  do ir = 1, nr
    do ib = 1, rotor(ir)%nb
      rotor(ir)%blade(ib)%theta = rotor(ir)%gettheta(rotor(ir)%psiStart, ib)
      call rotor(ir)%blade(ib)%rot_pitch(sign(1._dp, rotor(ir)%Omega)*rotor(ir)%blade(ib)%theta)
    enddo
    call rotor(ir)%calcAIC()
    rotor(ir)%omegaSlow = rotor(ir)%Omega
  enddo

  call solve_gamma(.false.)
  do ir = 1, nr
    rotor(ir)%rowFar = rotor(ir)%nFwake + 1
    rotor(ir)%rowNear = rotor(ir)%nNwake + 1
    if (rotor(ir)%nNwake > 0) call rotor(ir)%assignshed('TE')
  enddo

  open (newunit=u, file='Results/refprobe.txt', status='replace', action='write')
  write (u, '(I0)') nr
  do ir = 1, nr
    associate (r => rotor(ir))
      k = 0
      if (sign(1._dp, r%Omega*r%controlPitch(1)) > eps) k = 1
      write (u, '(10(I0,1X))') r%nb, r%nc, r%ns, r%nNwake, r%nFwake, r%nbConvect, &
        & r%rollupStart, r%rollupEnd, k, r%suppressFwakeSwitch
      write (u, '(ES26.17E3)') r%spanwiseCore
      write (u, '(ES26.17E3)') r%streamwiseCoreVec
      write (u, '(ES26.17E3)') r%apparentViscCoeff, r%decayCoeff
    end associate
  enddo
  write (u, '(I0)') nt
  write (u, '(ES26.17E3)') dt, kinematicVisc
  write (u, '(I0,1X,L1,1X,L1)') 0, switches%wakeDissipation == 1, switches%wakeStrain == 1
  call dump(u, 0)

  do iter = 1, nt
    do ir = 1, nr
      rotor(ir)%rowNear = max(rotor(ir)%rowNear - 1, 1)
      if (iter > rotor(ir)%nNwake) rotor(ir)%rowFar = max(rotor(ir)%rowFar - 1, 1)
      rotor(ir)%omegaSlow = rotor(ir)%Omega
    enddo
    do ir = 1, nr
      call rotor(ir)%move(rotor(ir)%velBody*dt)
      call rotor(ir)%rot_pts(rotor(ir)%omegaBody*dt, rotor(ir)%cgCoords, 1)
      call rotor(ir)%rot_advance(rotor(ir)%omegaSlow*dt)
    enddo
    do ir = 1, nr
      if (rotor(ir)%nNwake > 0) then
        call rotor(ir)%assignshed('LE')
        call rotor(ir)%age_wake(dt)
        if (switches%wakeDissipation == 1) call rotor(ir)%dissipate_wake(dt, kinematicVisc)
      endif
    enddo
    call solve_gamma(.true.)
    call convect_am2(iter)
    do ir = 1, nr
      if (rotor(ir)%nNwake > 0) then
        if (switches%wakeStrain == 1) call rotor(ir)%strain_wake()
      endif
    enddo
    do ir = 1, nr
      if (rotor(ir)%nNwake > 0) then
        if (rotor(ir)%rowNear == 1) call rotor(ir)%rollup()
        call rotor(ir)%assignshed('TE')
      endif
    enddo
    call dump(u, iter)
  enddo
  close (u)

contains

  subroutine solve_gamma(with_wake)
    !! RHS from the kinematic, wake-induced and other-wing velocity at every
    !! collocation point, gamVec = AIC_inv RHS, map_gam.
    logical, intent(in) :: with_wake
    integer :: irr, jrr, ibb, iss, icc, rw
    do irr = 1, nr
      associate (r => rotor(irr))
        r%RHS = 0._dp
        do ibb = 1, r%nbConvect
          do iss = 1, r%ns
            do icc = 1, r%nc
              rw = icc + r%nc*(iss - 1) + r%ns*r%nc*(ibb - 1)
              associate (w => r%blade(ibb)%wiP(icc, iss))
                w%velCP = -1._dp*r%velBody &
                  & - cross_product(r%omegaBody, w%CP - r%cgCoords) &
                  & - cross_product(r%omegaSlow*r%shaftAxis, w%CP - r%hubCoords) &
                  & - r%blade(ibb)%secMFlapArm(iss)*r%blade(ibb)%dflap
                do jrr = 1, nr
                  if (with_wake) w%velCP = w%velCP + rotor(jrr)%vind_bywake(w%CP)
                  if (irr /= jrr) w%velCP = w%velCP + rotor(jrr)%vind_bywing(w%CP)
                enddo
                r%RHS(rw) = dot_product(w%velCP, w%nCap)
              end associate
            enddo
          enddo
        enddo
        if (r%axisymmetrySwitch == 1) then
          do ibb = 2, r%nb
            r%RHS(r%nc*r%ns*(ibb - 1) + 1:r%nc*r%ns*ibb) = r%RHS(1:r%nc*r%ns)
          enddo
        endif
        r%RHS = -1._dp*r%RHS
      end associate
    enddo
    do irr = 1, nr
      rotor(irr)%gamVecPrev = rotor(irr)%gamVec
      rotor(irr)%gamVec = matmulAX(rotor(irr)%AIC_inv, rotor(irr)%RHS)
      call rotor(irr)%map_gam()
    enddo
  end subroutine solve_gamma

  subroutine wake_pass(predicted)
    !! Induced velocity of every rotor on the (predicted) wake nodes of the
    !! blades ib <= nbConvect, into velNwake(Predicted)/velFwake(Predicted).
    logical, intent(in) :: predicted
    integer :: irr, jrr, ibb, rn, ne, rf, fe
    do irr = 1, nr
      associate (r => rotor(irr))
        if (r%nNwake <= 0) cycle
        rn = r%rowNear; ne = r%nNwakeEnd; rf = r%rowFar; fe = r%nFwakeEnd
        do ibb = 1, r%nbConvect
          associate (b => r%blade(ibb))
            if (predicted) then
              b%velNwakePredicted(:, rn:ne, :) = 0._dp
              b%velFwakePredicted(:, rf:fe) = 0._dp
              do jrr = 1, nr
                b%velNwakePredicted(:, rn:ne, :) = b%velNwakePredicted(:, rn:ne, :) + &
                  & vind_onNwake_byRotor(rotor(jrr), b%waNPredicted(rn:ne, :), 'P')
                b%velFwakePredicted(:, rf:fe) = b%velFwakePredicted(:, rf:fe) + &
                  & vind_onFwake_byRotor(rotor(jrr), b%waFPredicted(rf:fe), 'P')
              enddo
            else
              b%velNwake(:, rn:r%nNwake, :) = 0._dp
              b%velFwake(:, rf:r%nFwake) = 0._dp
              do jrr = 1, nr
                b%velNwake(:, rn:ne, :) = b%velNwake(:, rn:ne, :) + &
                  & vind_onNwake_byRotor(rotor(jrr), b%waN(rn:ne, :))
                b%velFwake(:, rf:fe) = b%velFwake(:, rf:fe) + &
                  & vind_onFwake_byRotor(rotor(jrr), b%waF(rf:fe))
              enddo
            endif
          end associate
        enddo
      end associate
    enddo
  end subroutine wake_pass

  subroutine convect_am2(it)
    !! Pass 1, then the AM2 predictor-corrector (main.f90:1002-1115).
    integer, intent(in) :: it
    integer :: irr, ibb
    call wake_pass(.false.)
    if (it == 1) then
      do irr = 1, nr
        if (rotor(irr)%nNwake <= 0) cycle
        call rotor(irr)%convectwake(it, dt, 'C')
        do ibb = 1, rotor(irr)%nbConvect
          rotor(irr)%blade(ibb)%velNwake1 = rotor(irr)%blade(ibb)%velNwake
          rotor(irr)%blade(ibb)%velFwake1 = rotor(irr)%blade(ibb)%velFwake
        enddo
      enddo
      return
    endif
    do irr = 1, nr
      associate (r => rotor(irr))
        if (r%nNwake <= 0) cycle
        do ibb = 1, r%nbConvect
          associate (b => r%blade(ibb))
            b%waNPredicted(r%rowNear:r%nNwakeEnd, :) = b%waN(r%rowNear:r%nNwakeEnd, :)
            b%velNwakeStep = b%velNwake
            b%velNwake = 0.5_dp*(3._dp*b%velNwake - b%velNwake1)
            b%waFPredicted(r%rowFar:r%nFwakeEnd) = b%waF(r%rowFar:r%nFwakeEnd)
            b%velFwakeStep = b%velFwake
            b%velFwake = 0.5_dp*(3._dp*b%velFwake - b%velFwake1)
          end associate
        enddo
        call r%convectwake(it, dt, 'P')
      end associate
    enddo
    call wake_pass(.true.)
    do irr = 1, nr
      associate (r => rotor(irr))
        if (r%nNwake <= 0) cycle
        do ibb = 1, r%nbConvect
          r%blade(ibb)%velNwake = (r%blade(ibb)%velNwakePredicted + r%blade(ibb)%velNwakeStep)*0.5_dp
          r%blade(ibb)%velFwake = (r%blade(ibb)%velFwakePredicted + r%blade(ibb)%velFwakeStep)*0.5_dp
        enddo
        call r%convectwake(it, dt, 'C')
        do ibb = 1, r%nbConvect
          r%blade(ibb)%velNwake1 = r%blade(ibb)%velNwakeStep
          r%blade(ibb)%velFwake1 = r%blade(ibb)%velFwakeStep
        enddo
      end associate
    enddo
  end subroutine convect_am2

  subroutine dump(u, it)
    !! Per rotor: cursors, shaft and hub; per blade: wing rings (corners
    !! fc(:,1) of the four filaments, their core radii, gam), near-wake rings
    !! (same), far filaments (both ends, gam, rVc, rVc0, l0, lc), velNwake,
    !! velNwake1, velFwake, velFwake1.
    integer, intent(in) :: u, it
    integer :: irr, ibb, i, j, f
    write (u, '(A,1X,I0)') 'STEP', it
    do irr = 1, nr
      associate (r => rotor(irr))
        write (u, '(I0,1X,I0)') r%rowNear, r%rowFar
        write (u, '(ES26.17E3)') r%shaftAxis, r%hubCoords
        do ibb = 1, r%nb
          associate (b => r%blade(ibb))
            do j = 1, r%ns
              do i = 1, r%nc
                write (u, '(ES26.17E3)') ((b%wiP(i, j)%vr%vf(f)%fc(k, 1), k=1, 3), f=1, 4)
                write (u, '(ES26.17E3)') (b%wiP(i, j)%vr%vf(f)%rVc, f=1, 4)
                write (u, '(ES26.17E3)') b%wiP(i, j)%vr%gam
              enddo
            enddo
            do i = 1, r%nNwake
              do j = 1, r%ns
                write (u, '(ES26.17E3)') ((b%waN(i, j)%vr%vf(f)%fc(k, 1), k=1, 3), f=1, 4)
                write (u, '(ES26.17E3)') (b%waN(i, j)%vr%vf(f)%rVc, f=1, 4)
                write (u, '(ES26.17E3)') b%waN(i, j)%vr%gam
              enddo
            enddo
            do i = 1, r%nFwake
              write (u, '(ES26.17E3)') b%waF(i)%vf%fc(:, 1), b%waF(i)%vf%fc(:, 2)
              write (u, '(ES26.17E3)') b%waF(i)%gam, b%waF(i)%vf%rVc, b%waF(i)%vf%rVc0, &
                & b%waF(i)%vf%l0, b%waF(i)%vf%lc
            enddo
            if (ibb <= r%nbConvect) then
              write (u, '(ES26.17E3)') b%velNwake, b%velNwake1
              write (u, '(ES26.17E3)') b%velFwake, b%velFwake1
            endif
          end associate
        enddo
      end associate
    enddo
  end subroutine dump

end program refprobe
