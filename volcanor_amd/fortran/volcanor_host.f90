!! Host side of the MI355X free-wake solver, in Fortran.
!!
!! Keeps VOLCANOR's object model (rotor / blade / wing panel / vortex ring)
!! and its `.case` inputs (config.nml v0.5, geomXX.nml v0.15) as the
!! drop-in API.  Geometry, rigid motion and loads are O(panels) host work in
!! FP64; every Biot-Savart sum, the wake state and its updates, and the
!! circulation solve run on the GPU through libvlm (module vlm_c).
!!
!! Reference routines followed (cibinjoseph/VOLCANOR):
!!   readConfig            src/libCommon.f90:51-108
!!   rotor_readGeom        src/classdef.f90:2541-2767
!!   rotor_init            src/classdef.f90:2769-3891 (flat and PLOT3D blades)
!!   motion                src/classdef.f90:1092-1281, 4202-4291
!!   RHS / solve / loads   src/main.f90:119-234, 400-764
!!   blade_calc_force      src/classdef.f90:1704-1896 (+ 2058-2380 helpers)
!!   force2file            src/libPostprocess.f90:814-882
module volcanor_host
  use iso_c_binding
  use vlm_c
  implicit none
  private

  integer, parameter, public :: dp = kind(1.d0)
  real(dp), parameter :: pi = 4._dp*atan(1._dp)
  real(dp), parameter :: twoPi = 2._dp*pi
  real(dp), parameter :: eps = epsilon(1._dp)
  real(dp), parameter :: degToRad = pi/180._dp
  real(dp), parameter :: radToDeg = 180._dp/pi
  real(dp), parameter :: xAxis(3) = [1._dp, 0._dp, 0._dp]
  real(dp), parameter :: yAxis(3) = [0._dp, 1._dp, 0._dp]
  real(dp), parameter :: zAxis(3) = [0._dp, 0._dp, 1._dp]

  type, public :: switches_class
    integer :: ntSub = 0, ntSubInit = 0, wakeDissipation = 0, wakeStrain = 0
    integer :: wakeBurst = 0, wakeSuppress = 0, slowStart = 0, slowStartNt = 0
    integer :: wakeTipPlot = 0, wakePlot = 0, gridPlot = 0, rotorForcePlot = 0
    integer :: fdScheme = 3, probe = 0, initWakeVelNt = 0
    integer :: restartFromNt = 0, restartWriteNt = 0
  end type switches_class

  !! Vortex ring of a wing panel: corners c1..c4 and the core radius of its
  !! four filaments vf1 c1->c2, vf2 c2->c3, vf3 c3->c4, vf4 c4->c1.
  type, public :: vr_class
    real(dp) :: fc(3, 4) = 0._dp
    real(dp) :: rVc(4) = 0._dp
    real(dp) :: gam = 0._dp
  end type vr_class

  type, public :: wingpanel_class
    type(vr_class) :: vr
    real(dp) :: PC(3, 4) = 0._dp
    real(dp) :: CP(3) = 0._dp, nCap(3) = 0._dp
    real(dp) :: tauCapChord(3) = 0._dp, tauCapSpan(3) = 0._dp
    real(dp) :: velCP(3) = 0._dp, velCPm(3) = 0._dp, velCPTotal(3) = 0._dp
    real(dp) :: chordwiseResVel(3) = 0._dp
    real(dp) :: normalForce(3) = 0._dp, normalForceUnsteady(3) = 0._dp
    real(dp) :: gamPrev = 0._dp, gamTrapz = 0._dp
    real(dp) :: delP = 0._dp, delPUnsteady = 0._dp
    real(dp) :: panelArea = 0._dp, meanChord = 0._dp, meanSpan = 0._dp
  end type wingpanel_class

  type, public :: blade_class
    type(wingpanel_class), allocatable :: wiP(:, :)
    real(dp) :: xAxis(3), yAxis(3), zAxis(3)
    real(dp) :: xAxisAzi(3), yAxisAzi(3), zAxisAzi(3)
    real(dp) :: xAxisAziFlap(3), yAxisAziFlap(3), zAxisAziFlap(3)
    real(dp) :: theta = 0._dp, psi = 0._dp, pivotLE = 0._dp
    real(dp) :: flapOrigin(3) = 0._dp
    real(dp) :: lift(3) = 0._dp, drag(3) = 0._dp, forceInertial(3) = 0._dp
    real(dp) :: liftUnsteady(3) = 0._dp
    real(dp), allocatable :: secCP(:, :), secTauCapChord(:, :), secTauCapSpan(:, :)
    real(dp), allocatable :: secNormalVec(:, :), secChordwiseResVel(:, :)
    real(dp), allocatable :: secLift(:, :), secLiftDir(:, :), secDragDir(:, :)
    real(dp), allocatable :: secForceInertial(:, :), secLiftUnsteady(:, :)
    real(dp), allocatable :: secLiftInPlane(:, :), secLiftOutPlane(:, :)
    real(dp), allocatable :: secChord(:), secArea(:), secCL(:), secCLu(:)
    real(dp), allocatable :: secAlpha(:), secMflapArm(:)
    real(dp), allocatable :: secPhi(:), secTheta(:), secViz(:), secVix(:)
  end type blade_class

  type, public :: rotor_class
    character(len=2) :: id = '01'
    integer :: nb = 1, nc = 1, ns = 1, nNwake = 0, nFwake = 0, nbConvect = 1
    integer :: surfaceType = 0, propConvention = 0, spanSpacing = 1, chordSpacing = 1
    integer :: axisymmetrySwitch = 0, symmetricTau = 0, forceCalcSwitch = 0
    integer :: wakeTruncateNt = 0, rollupStart = 1, rollupEnd = 1
    integer :: prescWakeAfterTruncNt = 0, prescWakeGenNt = 0, prescWakeNt = 0
    integer :: spanwiseLiftSwitch = 0, suppressFwakeSwitch = 0
    integer :: rowNear = 0, rowFar = 0
    character(len=30) :: geometryFile = '0'
    type(blade_class), allocatable :: blade(:)
    real(dp) :: Omega = 0._dp, omegaSlow = 0._dp, psi = 0._dp
    real(dp) :: shaftAxis(3) = zAxis, hubCoords(3) = 0._dp, cgCoords(3) = 0._dp
    real(dp) :: fromCoords(3) = 0._dp, pts(3) = 0._dp
    real(dp) :: xAxisBody(3) = xAxis, yAxisBody(3) = yAxis, zAxisBody(3) = zAxis
    real(dp) :: radius = 1._dp, root_cut = 0._dp, chord = 1._dp
    real(dp) :: preconeAngle = 0._dp, controlPitch(3) = 0._dp, pivotLE = 0._dp
    real(dp) :: flapHinge = 0._dp, psiStart = 0._dp
    real(dp) :: velBody(3) = 0._dp, omegaBody(3) = 0._dp
    real(dp) :: spanwiseCore = 0._dp, apparentViscCoeff = 0._dp, decayCoeff = 0._dp
    real(dp) :: rollupStartRadius = 0._dp, rollupEndRadius = 1._dp
    real(dp), allocatable :: streamwiseCoreVec(:)
    real(dp), allocatable :: gamVec(:), gamVecPrev(:), RHS(:)
    real(dp), allocatable :: lout(:)  !! packed loads of vlm_rotor_loads (deferred download)
    ! inputs kept only for rXXParams.json (libPostprocess.f90:52-141)
    integer :: imagePlane = 0, imageRotorNum = 0
    real(dp) :: thetaTwist = 0._dp, initWakeVel = 0._dp
    integer :: inflowPlotSwitch = 0, gammaPlotSwitch = 0, customTrajectorySwitch = 0
    integer :: nCamberFiles = 0, skewPlotSwitch = 0
    real(dp) :: camberSectionLimit(8) = 0._dp
    character(len=30) :: camberFile(8) = '0'
    real(dp) :: skewLimit = 0._dp
    real(dp), allocatable :: velBodyHistory(:, :), omegaBodyHistory(:, :)
    real(dp), allocatable :: inflowV(:, :, :)
    ! azimuthal age of the near-wake rows and far-wake slots, for the tip
    ! plot: the same in every column, filament and blade, so one value per row
    real(dp), allocatable :: ageAzN(:), ageAzF(:)
    real(dp) :: flapInitial = 0._dp, dflapInitial = 0._dp, Iflap = 0._dp, cflap = 0._dp
    real(dp) :: kflap = 0._dp, MflapConstant = 0._dp, dpitch = 0._dp
    integer(c_int) :: solve_iters = 0
    real(dp) :: lift(3) = 0._dp, drag(3) = 0._dp, forceInertial(3) = 0._dp
    real(dp) :: liftUnsteady(3) = 0._dp
    real(dp) :: nonDimforceDenominator = 1._dp
    integer(c_int) :: dev = -1        !! libvlm rotor index
  end type rotor_class

  !! Result of one run, returned to C callers.
  type, bind(C), public :: volcanor_stats
    real(c_double) :: wall_total, wall_loop, wall_steps_tail
    real(c_double) :: bs_kernel_ms, bs_pairs
    integer(c_int64_t) :: bs_launches
    integer(c_int) :: nt, tail_steps
    integer(c_int64_t) :: pred_full, pred_learned, pred_fallbacks
    real(c_double) :: pred_last_err
    real(c_double) :: pred_mlp_rms, pred_base_rms
    integer(c_int64_t) :: pred_judged
    real(c_double) :: bs_pairs_all         !! every Biot-Savart launch of the tail
    integer(c_int64_t) :: bs_launches_all
  end type volcanor_stats

  public :: volcanor_run_impl, volcanor_host_cleanup, run_case

  !! libvlm context of the run in progress (destroyed by
  !! volcanor_host_cleanup when an error stops the run)
  type(c_ptr), save :: active_ctx = c_null_ptr

  interface
    integer(c_int) function c_mkdir(path, mode) bind(C, name='mkdir')
      !! POSIX mkdir(2); an existing directory is not an error here
      import :: c_int, c_char
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), value :: mode
    end function c_mkdir
  end interface

contains

  ! ------------------------------------------------------------------
  ! Vector helpers (own implementations of the libMath routines used)
  ! ------------------------------------------------------------------
  pure function cross3(a, b) result(c)
    real(dp), intent(in) :: a(3), b(3)
    real(dp) :: c(3)
    c(1) = a(2)*b(3) - a(3)*b(2)
    c(2) = a(3)*b(1) - a(1)*b(3)
    c(3) = a(1)*b(2) - a(2)*b(1)
  end function cross3

  pure function unit3(a) result(u)
    real(dp), intent(in) :: a(3)
    real(dp) :: u(3), n
    n = norm2(a)
    if (n > eps) then
      u = a/n
    else
      u = 0._dp
    endif
  end function unit3

  pure function proj3(a, d) result(p)
    real(dp), intent(in) :: a(3), d(3)
    real(dp) :: p(3), n2
    n2 = sum(d**2)
    if (n2 > eps) then
      p = dot_product(a, d)*d/n2
    else
      p = 0._dp
    endif
  end function proj3

  function spaced(kind, x0, x1, n) result(x)
    !! linspace / cosspace / halfsinspace / tanspace, libMath.f90:138-200
    integer, intent(in) :: kind, n
    real(dp), intent(in) :: x0, x1
    real(dp) :: x(n), t(n), dx
    integer :: i
    select case (kind)
    case (1)
      dx = (x1 - x0)/(n - 1)
      do i = 1, n
        x(i) = (i - 1)*dx + x0
      enddo
    case (2)
      dx = pi/(n - 1)
      do i = 1, n
        t(i) = (i - 1)*dx
      enddo
      x = x0 + (x1 - x0)*0.5_dp*(1._dp - cos(t))
    case (3)
      dx = (0.5_dp*pi)/(n - 1)
      do i = 1, n
        t(i) = (i - 1)*dx
      enddo
      x = x0 + (x1 - x0)*sin(t)
    case (4)
      dx = 2.4_dp/(n - 1)
      do i = 1, n
        t(i) = (i - 1)*dx - 1.2_dp
      enddo
      x = x0 + (x1 - x0)*tan(t)/tan(1.2_dp)
    case default
      call host_fail('ERROR: spacing must be 1..4')
    end select
  end function spaced

  function axis_rotation(theta, ax) result(T)
    !! Rotation by theta about a unit axis, column-major, x' = T x
    real(dp), intent(in) :: theta, ax(3)
    real(dp) :: T(3, 3), a(3), ct, st, om
    a = ax/norm2(ax)
    ct = cos(theta)
    st = sin(theta)
    om = 1._dp - ct
    T(:, 1) = [ct + a(1)*a(1)*om, a(3)*st + a(2)*a(1)*om, -a(2)*st + a(3)*a(1)*om]
    T(:, 2) = [-a(3)*st + a(1)*a(2)*om, ct + a(2)*a(2)*om, a(1)*st + a(3)*a(2)*om]
    T(:, 3) = [a(2)*st + a(1)*a(3)*om, -a(1)*st + a(2)*a(3)*om, ct + a(3)*a(3)*om]
  end function axis_rotation

  function euler_gb(p) result(T)
    !! Global-to-body Euler matrix from (phi, theta, psi), libMath.f90:672-693
    real(dp), intent(in) :: p(3)
    real(dp) :: T(3, 3), cf, sf, ct, st, cs, ss
    cf = cos(p(1)); sf = sin(p(1))
    ct = cos(p(2)); st = sin(p(2))
    cs = cos(p(3)); ss = sin(p(3))
    T(1, 1) = cs*ct
    T(1, 2) = sf*st*cs - ss*cf
    T(1, 3) = sf*ss + st*cf*cs
    T(2, 1) = ss*ct
    T(2, 2) = sf*ss*st + cf*cs
    T(2, 3) = ss*st*cf - sf*cs
    T(3, 1) = -st
    T(3, 2) = sf*ct
    T(3, 3) = cf*ct
  end function euler_gb

  function wing_camber(r, dir, x, y) result(z)
    !! rotor_getCamber (classdef.f90:4060-4112): z of the wing grid from the
    !! camber lines (files of <n> then n lines "x z", normalised by the last
    !! x; '0' is a flat line).  Node column j uses the first file whose
    !! sectionLimit is >= y(j)/span (else the previous column's file), and
    !! z = chord * interp1d((x - x(1))/chord) of order 2, with chord and span
    !! the signed extents of x and y, as the reference computes them.
    type(rotor_class), intent(in) :: r
    character(len=*), intent(in) :: dir
    real(dp), intent(in) :: x(:), y(:)
    real(dp) :: z(size(x), size(y)), chord, span
    real(dp), allocatable :: xc(:, :), zc(:, :)
    integer :: np(r%nCamberFiles), j, i, k, u, f
    allocate (xc(5000, r%nCamberFiles), zc(5000, r%nCamberFiles))
    chord = x(size(x)) - x(1)
    span = y(size(y)) - y(1)
    do k = 1, r%nCamberFiles
      if (r%camberFile(k)(1:1) == '0') then
        np(k) = 2
        xc(1:2, k) = [0._dp, 1._dp]
        zc(1:2, k) = 0._dp
      else
        call open_old(u, trim(dir)//'/'//trim(r%camberFile(k)))
        read (u, *) np(k)
        do i = 1, np(k)
          read (u, *) xc(i, k), zc(i, k)
        enddo
        close (u)
        zc(:, k) = zc(:, k)/xc(np(k), k)
        xc(:, k) = xc(:, k)/xc(np(k), k)
      endif
    enddo
    f = 1
    do j = 1, size(y)
      do k = 1, r%nCamberFiles
        if (y(j)/span <= r%camberSectionLimit(k)) then
          f = k
          exit
        endif
      enddo
      do i = 1, size(x)
        z(i, j) = chord*lagrange2((x(i) - x(1))/chord, xc(1:np(f), f), zc(1:np(f), f))
      enddo
    enddo
  end function wing_camber

  pure function lagrange2(xq, x, y) result(v)
    !! interp1d of order 2 (libMath.f90:520-575): the value at the first node
    !! with x >= xq when xq is within eps of it, 0 when both bracketing
    !! values vanish, else the quadratic Lagrange polynomial through that
    !! node, its predecessor and, on the first interval, its successor or,
    !! after it, the node before the predecessor (linear with two nodes)
    real(dp), intent(in) :: xq, x(:), y(:)
    real(dp) :: v, l0, l1, l2
    integer :: ix, i
    ix = size(x)
    do i = 1, size(x)
      if (xq <= x(i)) then
        ix = i
        exit
      endif
    enddo
    if (abs(xq - x(ix)) <= eps) then
      v = y(ix)
    elseif (norm2(y(ix - 1:ix)) <= eps) then
      v = 0._dp
    elseif (size(x) == 2) then
      v = y(ix - 1) + (xq - x(ix - 1))*(y(ix) - y(ix - 1))/(x(ix) - x(ix - 1))
    elseif (ix == 2) then
      l0 = (xq - x(ix))*(xq - x(ix + 1))/((x(ix - 1) - x(ix))*(x(ix - 1) - x(ix + 1)))
      l1 = (xq - x(ix - 1))*(xq - x(ix + 1))/((x(ix) - x(ix - 1))*(x(ix) - x(ix + 1)))
      l2 = (xq - x(ix - 1))*(xq - x(ix))/((x(ix + 1) - x(ix - 1))*(x(ix + 1) - x(ix)))
      v = y(ix - 1)*l0 + y(ix)*l1 + y(ix + 1)*l2
    else
      l0 = (xq - x(ix - 1))*(xq - x(ix))/((x(ix - 2) - x(ix - 1))*(x(ix - 2) - x(ix)))
      l1 = (xq - x(ix - 2))*(xq - x(ix))/((x(ix - 1) - x(ix - 2))*(x(ix - 1) - x(ix)))
      l2 = (xq - x(ix - 2))*(xq - x(ix - 1))/((x(ix) - x(ix - 2))*(x(ix) - x(ix - 1)))
      v = y(ix - 2)*l0 + y(ix - 1)*l1 + y(ix)*l2
    endif
  end function lagrange2

  pure function slow_omega(sw, iter, Omega) result(w)
    !! Rotor speed of the slow start, main.f90:428-452 (the reference's
    !! default-real conversions kept; they are exact for step counts)
    type(switches_class), intent(in) :: sw
    integer, intent(in) :: iter
    real(dp), intent(in) :: Omega
    real(dp) :: w
    select case (sw%slowStart)
    case (1)
      w = min(real(sw%slowStartNt), real(iter + 1))*Omega/sw%slowStartNt
    case (2)
      w = tanh(5._dp*iter/sw%slowStartNt)*Omega
    case (3)
      w = (tanh(6._dp*real(iter)/real(sw%slowStartNt) - 3._dp) + 1._dp)*0.5_dp*Omega
    case default
      w = Omega
    end select
  end function slow_omega

  function pwl(x, y, q) result(v)
    !! Piecewise-linear interpolation, libMath.f90:476-517
    real(dp), intent(in) :: x(:), y(:), q
    real(dp) :: v
    integer :: n, i, idx
    n = size(x)
    if (abs(x(1) - q) < eps) then
      v = y(1)
      return
    elseif (abs(x(n) - q) < eps) then
      v = y(n)
      return
    endif
    idx = 1
    do i = 1, n
      if (x(1) < x(n)) then
        if (.not. (x(i) <= q)) then
          idx = i - 1
          exit
        endif
      else
        if (.not. (x(i) >= q)) then
          idx = i - 1
          exit
        endif
      endif
    enddo
    if (idx < 1) call host_fail('ERROR: interpolation point out of range')
    v = y(idx) + (y(idx + 1) - y(idx))/(x(idx + 1) - x(idx))*(q - x(idx))
  end function pwl

  ! ------------------------------------------------------------------
  ! Inputs
  ! ------------------------------------------------------------------
  subroutine open_old(u, path, unformatted)
    !! Opens an existing input file for reading; a missing or unreadable
    !! file stops the run with the runtime's message.
    integer, intent(out) :: u
    character(len=*), intent(in) :: path
    logical, intent(in), optional :: unformatted
    integer :: ios
    character(len=256) :: emsg
    logical :: unf
    unf = .false.
    if (present(unformatted)) unf = unformatted
    emsg = ''
    if (unf) then
      open (newunit=u, file=path, status='old', action='read', form='unformatted', &
        & iostat=ios, iomsg=emsg)
    else
      open (newunit=u, file=path, status='old', action='read', iostat=ios, iomsg=emsg)
    endif
    call io_check(ios, emsg, path)
  end subroutine open_old

  subroutine io_check(ios, emsg, what)
    !! Stops the run with the runtime's message when an input open or
    !! namelist read failed.
    integer, intent(in) :: ios
    character(len=*), intent(in) :: emsg, what
    if (ios /= 0) call host_fail('ERROR: reading '//what//': '//trim(emsg))
  end subroutine io_check

  subroutine read_config(dir, nt, dt, nr, density, velSound, kinematicVisc, sw)
    !! config.nml v0.5 (libCommon.f90:51-108); absent keys default to 0
    character(len=*), intent(in) :: dir
    integer, intent(out) :: nt, nr
    real(dp), intent(out) :: dt, density, velSound, kinematicVisc
    type(switches_class), intent(out) :: sw
    character(len=10) :: fileFormatVersion
    integer :: restartWriteNt, restartFromNt, ntSub, ntSubInit, wakePlot, wakeTipPlot, &
      & rotorForcePlot, gridPlot, wakeDissipation, wakeStrain, wakeBurst, wakeSuppress, &
      & slowStart, slowStartNt, fdScheme, initWakeVelNt, probe
    integer :: u, ios
    character(len=256) :: emsg
    namelist /VERSION/ fileFormatVersion
    namelist /PARAMS/ nt, dt, nr, density, velSound, kinematicVisc
    namelist /OPTIONS/ restartWriteNt, restartFromNt, ntSub, ntSubInit, &
      & wakePlot, wakeTipPlot, rotorForcePlot, gridPlot, wakeDissipation, wakeStrain, &
      & wakeBurst, wakeSuppress, slowStart, slowStartNt, fdScheme, initWakeVelNt, probe

    nt = 0; dt = 0._dp; nr = 1; density = 1._dp; velSound = 330._dp; kinematicVisc = 0._dp
    restartWriteNt = 0; restartFromNt = 0; ntSub = 0; ntSubInit = 0; wakePlot = 0
    wakeTipPlot = 0; rotorForcePlot = 0; gridPlot = 0; wakeDissipation = 0; wakeStrain = 0
    wakeBurst = 0; wakeSuppress = 0; slowStart = 0; slowStartNt = 0; fdScheme = 3
    initWakeVelNt = 0; probe = 0
    open (newunit=u, file=trim(dir)//'/config.nml', status='old', action='read', iostat=ios, &
      & iomsg=emsg)
    call io_check(ios, emsg, trim(dir)//'/config.nml')
    read (u, nml=VERSION, iostat=ios, iomsg=emsg)
    call io_check(ios, emsg, 'config.nml &VERSION')
    if (adjustl(fileFormatVersion) /= '0.5') &
      call host_fail('ERROR: config.nml template version does not match')
    read (u, nml=PARAMS, iostat=ios, iomsg=emsg)
    call io_check(ios, emsg, 'config.nml &PARAMS')
    read (u, nml=OPTIONS, iostat=ios, iomsg=emsg)
    call io_check(ios, emsg, 'config.nml &OPTIONS')
    close (u)
    sw%restartWriteNt = restartWriteNt; sw%restartFromNt = restartFromNt
    sw%ntSub = ntSub; sw%ntSubInit = ntSubInit; sw%wakePlot = wakePlot
    sw%wakeTipPlot = wakeTipPlot; sw%rotorForcePlot = rotorForcePlot; sw%gridPlot = gridPlot
    sw%wakeDissipation = wakeDissipation; sw%wakeStrain = wakeStrain
    sw%wakeBurst = wakeBurst; sw%wakeSuppress = wakeSuppress; sw%slowStart = slowStart
    sw%slowStartNt = slowStartNt; sw%fdScheme = fdScheme; sw%initWakeVelNt = initWakeVelNt
    sw%probe = probe
  end subroutine read_config

  subroutine rotor_readGeom(r, dir, fname)
    !! geomXX.nml v0.15 (classdef.f90:2541-2767); absent keys default to 0
    type(rotor_class), intent(inout) :: r
    character(len=*), intent(in) :: dir, fname
    character(len=10) :: fileFormatVersion
    integer :: surfaceType, imagePlane, imageRotorNum, nb, propConvention, spanSpacing, &
      & chordSpacing, nCamberFiles, nc, ns, nNwake, nAirfoils
    real(dp) :: camberSectionLimit(8)
    character(len=30) :: camberFile(8), geometryFile
    real(dp), dimension(3) :: hubCoords, cgCoords, fromCoords, phiThetaPsi, shaftAxis, &
      & velBody, omegaBody, dragUnitVec, sideUnitVec, liftUnitVec
    real(dp) :: span, rootcut, chord, preconeAngle, Omega, theta0, thetaC, thetaS, &
      & thetaTwist, pivotLE, flapHinge
    integer :: spanwiseLiftSwitch, symmetricTau, ductSwitch, axisymmetrySwitch, &
      & customTrajectorySwitch, forceCalcSwitch, wakeTruncateNt, prescWakeAfterTruncNt, &
      & prescWakeGenNt
    real(dp) :: apparentViscCoeff, decayCoeff, spanwiseCore, rollupStartRadius, &
      & rollupEndRadius, initWakeVel, psiStart, skewLimit
    real(dp), allocatable :: streamwiseCoreVec(:)
    integer :: bladeDynamicsSwitch, pitchDynamicsSwitch, bodyDynamicsSwitch, &
      & bodyDynamicsIOVars, inflowPlotSwitch, gammaPlotSwitch, skewPlotSwitch
    real(dp) :: flapInitial, dflapInitial, Iflap, cflap, kflap, MflapConstant, dpitch
    integer :: u, ios
    character(len=256) :: emsg
    namelist /VERSION/ fileFormatVersion
    namelist /SURFACE/ surfaceType, imagePlane, imageRotorNum
    namelist /PANELS/ nb, propConvention, spanSpacing, chordSpacing, geometryFile, &
      & nCamberFiles, nc, ns, nNwake
    namelist /CAMBERSECTIONS/ camberSectionLimit, camberFile
    namelist /ORIENT/ hubCoords, cgCoords, fromCoords, phiThetaPsi
    namelist /GEOMPARAMS/ span, rootcut, chord, preconeAngle, Omega, shaftAxis, theta0, &
      & thetaC, thetaS, thetaTwist, ductSwitch, axisymmetrySwitch, pivotLE, flapHinge, &
      & spanwiseLiftSwitch, symmetricTau, customTrajectorySwitch, velBody, omegaBody, &
      & forceCalcSwitch, nAirfoils
    namelist /WAKEPARAMS/ apparentViscCoeff, decayCoeff, wakeTruncateNt, &
      & prescWakeAfterTruncNt, prescWakeGenNt, spanwiseCore, streamwiseCoreVec, &
      & rollupStartRadius, rollupEndRadius, initWakeVel, psiStart, skewLimit
    namelist /DYNAMICS/ bladeDynamicsSwitch, flapInitial, dflapInitial, Iflap, cflap, &
      & kflap, MflapConstant, pitchDynamicsSwitch, dpitch, bodyDynamicsSwitch, &
      & bodyDynamicsIOVars
    namelist /WINDFRAME/ dragUnitVec, sideUnitVec, liftUnitVec
    namelist /PLOTS/ inflowPlotSwitch, gammaPlotSwitch, skewPlotSwitch

    surfaceType = 0; imagePlane = 0; imageRotorNum = 0; nb = 1; propConvention = 0
    spanSpacing = 1; chordSpacing = 1; nCamberFiles = 0; nc = 1; ns = 1; nNwake = 0
    nAirfoils = 0; camberSectionLimit = 0._dp; camberFile = '0'; geometryFile = '0'
    hubCoords = 0._dp; cgCoords = 0._dp; fromCoords = 0._dp; phiThetaPsi = 0._dp
    shaftAxis = zAxis; velBody = 0._dp; omegaBody = 0._dp
    dragUnitVec = 0._dp; sideUnitVec = 0._dp; liftUnitVec = 0._dp
    span = 1._dp; rootcut = 0._dp; chord = 1._dp; preconeAngle = 0._dp; Omega = 0._dp
    theta0 = 0._dp; thetaC = 0._dp; thetaS = 0._dp; thetaTwist = 0._dp; pivotLE = 0._dp
    flapHinge = 0._dp; spanwiseLiftSwitch = 0; symmetricTau = 0; ductSwitch = 0
    axisymmetrySwitch = 0; customTrajectorySwitch = 0; forceCalcSwitch = 0
    wakeTruncateNt = 0; prescWakeAfterTruncNt = 0; prescWakeGenNt = 0
    apparentViscCoeff = 0._dp; decayCoeff = 0._dp; spanwiseCore = 0._dp
    rollupStartRadius = 0._dp; rollupEndRadius = 1._dp; initWakeVel = 0._dp
    psiStart = 0._dp; skewLimit = 0._dp
    bladeDynamicsSwitch = 0; pitchDynamicsSwitch = 0; bodyDynamicsSwitch = 0
    bodyDynamicsIOVars = 0; inflowPlotSwitch = 0; gammaPlotSwitch = 0; skewPlotSwitch = 0
    flapInitial = 0._dp; dflapInitial = 0._dp; Iflap = 0._dp; cflap = 0._dp; kflap = 0._dp
    MflapConstant = 0._dp; dpitch = 0._dp

    open (newunit=u, file=trim(dir)//'/'//trim(fname), status='old', action='read', &
      & iostat=ios, iomsg=emsg)
    call io_check(ios, emsg, trim(dir)//'/'//trim(fname))
    read (u, nml=VERSION, iostat=ios, iomsg=emsg)
    call io_check(ios, emsg, trim(fname)//' &VERSION')
    if (adjustl(fileFormatVersion) /= '0.15') &
      call host_fail('ERROR: geomXX.nml template version does not match')
    read (u, nml=SURFACE, iostat=ios, iomsg=emsg)
    call io_check(ios, emsg, trim(fname)//' &SURFACE')
    read (u, nml=PANELS, iostat=ios, iomsg=emsg)
    call io_check(ios, emsg, trim(fname)//' &PANELS')
    read (u, nml=CAMBERSECTIONS, iostat=ios, iomsg=emsg)
    call io_check(ios, emsg, trim(fname)//' &CAMBERSECTIONS')
    read (u, nml=ORIENT, iostat=ios, iomsg=emsg)
    call io_check(ios, emsg, trim(fname)//' &ORIENT')
    read (u, nml=GEOMPARAMS, iostat=ios, iomsg=emsg)
    call io_check(ios, emsg, trim(fname)//' &GEOMPARAMS')
    allocate (streamwiseCoreVec(ns + 1))
    streamwiseCoreVec = 0._dp
    read (u, nml=WAKEPARAMS, iostat=ios, iomsg=emsg)
    call io_check(ios, emsg, trim(fname)//' &WAKEPARAMS')
    read (u, nml=DYNAMICS, iostat=ios, iomsg=emsg)
    call io_check(ios, emsg, trim(fname)//' &DYNAMICS')
    read (u, nml=WINDFRAME, iostat=ios, iomsg=emsg)
    call io_check(ios, emsg, trim(fname)//' &WINDFRAME')
    read (u, nml=PLOTS, iostat=ios, iomsg=emsg)
    call io_check(ios, emsg, trim(fname)//' &PLOTS')
    close (u)

    if (abs(surfaceType) == 2) call host_fail('ERROR: non-lifting surfaces are not supported')
    r%surfaceType = surfaceType
    r%imagePlane = imagePlane; r%imageRotorNum = imageRotorNum
    r%nb = nb; r%propConvention = propConvention; r%spanSpacing = spanSpacing
    r%chordSpacing = chordSpacing; r%geometryFile = geometryFile
    r%nc = nc; r%ns = ns; r%nNwake = nNwake
    if (surfaceType < 0) then
      ! an image keeps only the entries above; rotor_init takes the rest
      ! from its source (classdef.f90:2657-2675, 2798-2955)
      if (imagePlane < 1 .or. imagePlane > 3) call host_fail('ERROR: imagePlane must be 1, 2 or 3')
      return
    endif
    if (nCamberFiles > 8) call host_fail('ERROR: at most 8 camber files')
    r%nCamberFiles = nCamberFiles
    r%camberSectionLimit = camberSectionLimit; r%camberFile = camberFile
    if (ductSwitch /= 0) call host_fail('ERROR: ducts are not supported')
    if (forceCalcSwitch /= 0) call host_fail('ERROR: only forceCalcSwitch = 0 is supported')
    if (bladeDynamicsSwitch /= 0 .or. bodyDynamicsSwitch /= 0 .or. pitchDynamicsSwitch /= 0) &
      call host_fail('ERROR: blade, pitch and body dynamics are not supported')
    if (customTrajectorySwitch /= 0 .and. customTrajectorySwitch /= 1) &
      & call host_fail('ERROR: customTrajectorySwitch must be 0 or 1')
    r%customTrajectorySwitch = customTrajectorySwitch
    r%prescWakeAfterTruncNt = prescWakeAfterTruncNt; r%prescWakeGenNt = prescWakeGenNt

    r%hubCoords = hubCoords; r%cgCoords = cgCoords; r%fromCoords = fromCoords
    r%pts = phiThetaPsi
    r%radius = span; r%root_cut = rootcut; r%chord = chord; r%preconeAngle = preconeAngle
    r%Omega = Omega; r%shaftAxis = shaftAxis; r%controlPitch = [theta0, thetaC, thetaS]
    r%axisymmetrySwitch = axisymmetrySwitch; r%pivotLE = pivotLE; r%flapHinge = flapHinge
    r%spanwiseLiftSwitch = spanwiseLiftSwitch; r%symmetricTau = symmetricTau
    r%velBody = velBody; r%omegaBody = omegaBody; r%forceCalcSwitch = forceCalcSwitch
    r%apparentViscCoeff = apparentViscCoeff; r%decayCoeff = decayCoeff
    r%wakeTruncateNt = wakeTruncateNt; r%spanwiseCore = spanwiseCore
    allocate (r%streamwiseCoreVec(ns + 1))
    if (norm2(streamwiseCoreVec(2:)) < eps) then
      r%streamwiseCoreVec = streamwiseCoreVec(1)
    else
      r%streamwiseCoreVec = streamwiseCoreVec
    endif
    r%rollupStartRadius = rollupStartRadius; r%rollupEndRadius = rollupEndRadius
    r%psiStart = psiStart
    r%thetaTwist = thetaTwist; r%initWakeVel = initWakeVel
    r%inflowPlotSwitch = inflowPlotSwitch; r%gammaPlotSwitch = gammaPlotSwitch
    r%skewLimit = skewLimit
    r%skewPlotSwitch = skewPlotSwitch
    r%flapInitial = flapInitial; r%dflapInitial = dflapInitial; r%Iflap = Iflap
    r%cflap = cflap; r%kflap = kflap; r%MflapConstant = MflapConstant; r%dpitch = dpitch
  end subroutine rotor_readGeom

  subroutine to_chords_revs(r, n, dt)
    !! Negative counts mean chords or revolutions, classdef.f90:5133-5148
    type(rotor_class), intent(in) :: r
    integer, intent(inout) :: n
    real(dp), intent(in) :: dt
    if (n < 0) then
      if (abs(r%Omega) < eps) then
        n = ceiling(abs(n)*r%chord/(dt*norm2(r%velBody)))
      else
        n = ceiling(twoPi*abs(n)/(abs(r%Omega)*dt))
      endif
    endif
  end subroutine to_chords_revs

  ! ------------------------------------------------------------------
  ! Rigid motion of blades
  ! ------------------------------------------------------------------
  subroutine panel_shift(w, d)
    type(wingpanel_class), intent(inout) :: w
    real(dp), intent(in) :: d(3)
    integer :: k
    w%CP = w%CP + d
    do k = 1, 4
      w%PC(:, k) = w%PC(:, k) + d
      w%vr%fc(:, k) = w%vr%fc(:, k) + d
    enddo
  end subroutine panel_shift

  subroutine panel_rot(w, T, o)
    type(wingpanel_class), intent(inout) :: w
    real(dp), intent(in) :: T(3, 3), o(3)
    integer :: k
    do k = 1, 4
      w%PC(:, k) = matmul(T, w%PC(:, k) - o) + o
      w%vr%fc(:, k) = matmul(T, w%vr%fc(:, k) - o) + o
    enddo
    w%CP = matmul(T, w%CP - o) + o
    w%nCap = matmul(T, w%nCap)
    w%tauCapChord = matmul(T, w%tauCapChord)
    w%tauCapSpan = matmul(T, w%tauCapSpan)
  end subroutine panel_rot

  subroutine blade_move(b, d)
    !! classdef.f90:1092-1112
    type(blade_class), intent(inout) :: b
    real(dp), intent(in) :: d(3)
    integer :: i, j
    do j = 1, size(b%wiP, 2)
      do i = 1, size(b%wiP, 1)
        call panel_shift(b%wiP(i, j), d)
      enddo
      b%secCP(:, j) = b%secCP(:, j) + d
    enddo
    b%flapOrigin = b%flapOrigin + d
  end subroutine blade_move

  subroutine blade_rotate(b, ang, ax, org, kind)
    !! Rotation about an axis through org: translate, rotate, translate
    !! back (classdef.f90:1193-1281); kind 1 azimuth, 2 flap, 3 pitch
    type(blade_class), intent(inout) :: b
    real(dp), intent(in) :: ang, ax(3), org(3)
    integer, intent(in) :: kind
    real(dp) :: T(3, 3), zero(3)
    integer :: i, j
    if (abs(ang) <= eps) return
    zero = 0._dp
    call blade_move(b, -org)
    T = axis_rotation(ang, ax)
    do j = 1, size(b%wiP, 2)
      do i = 1, size(b%wiP, 1)
        call panel_rot(b%wiP(i, j), T, zero)
      enddo
    enddo
    call blade_move(b, org)
    do j = 1, size(b%wiP, 2)
      b%secCP(:, j) = matmul(T, b%secCP(:, j) - org) + org
      b%secTauCapChord(:, j) = matmul(T, b%secTauCapChord(:, j))
      b%secTauCapSpan(:, j) = matmul(T, b%secTauCapSpan(:, j))
      b%secNormalVec(:, j) = matmul(T, b%secNormalVec(:, j))
    enddo
    if (kind == 1 .or. kind == 2) then
      b%xAxisAziFlap = matmul(T, b%xAxisAziFlap)
      b%yAxisAziFlap = matmul(T, b%yAxisAziFlap)
      b%zAxisAziFlap = matmul(T, b%zAxisAziFlap)
    endif
    if (kind == 1) then
      b%xAxisAzi = matmul(T, b%xAxisAzi)
      b%yAxisAzi = matmul(T, b%yAxisAzi)
      b%zAxisAzi = matmul(T, b%zAxisAzi)
    endif
    b%xAxis = matmul(T, b%xAxis)
    b%yAxis = matmul(T, b%yAxis)
    b%zAxis = matmul(T, b%zAxis)
  end subroutine blade_rotate

  subroutine blade_rot_pitch(b, theta)
    !! Pitch about the line through the root LE/TE pivot, classdef.f90:1164-1181
    type(blade_class), intent(inout) :: b
    real(dp), intent(in) :: theta
    real(dp) :: org(3)
    integer :: nc
    if (abs(theta) <= eps) return
    nc = size(b%wiP, 1)
    org = b%wiP(1, 1)%PC(:, 1)*(1._dp - b%pivotLE) + b%wiP(nc, 1)%PC(:, 2)*b%pivotLE
    call blade_rotate(b, theta, b%yAxis, org, 3)
  end subroutine blade_rot_pitch

  subroutine rotor_rot_pts(r, p, org)
    !! Euler rotation of the whole rotor about org, classdef.f90:4202-4240
    type(rotor_class), intent(inout) :: r
    real(dp), intent(in) :: p(3), org(3)
    real(dp) :: T(3, 3)
    integer :: ib, i, j
    T = euler_gb(p)
    do ib = 1, r%nb
      associate (b => r%blade(ib))
        do j = 1, r%ns
          do i = 1, r%nc
            call panel_rot(b%wiP(i, j), T, org)
          enddo
          b%secCP(:, j) = matmul(T, b%secCP(:, j) - org) + org
          b%secTauCapChord(:, j) = matmul(T, b%secTauCapChord(:, j))
          b%secNormalVec(:, j) = matmul(T, b%secNormalVec(:, j))
        enddo
        b%xAxis = matmul(T, b%xAxis); b%yAxis = matmul(T, b%yAxis)
        b%zAxis = matmul(T, b%zAxis)
        b%xAxisAzi = matmul(T, b%xAxisAzi); b%yAxisAzi = matmul(T, b%yAxisAzi)
        b%zAxisAzi = matmul(T, b%zAxisAzi)
        b%xAxisAziFlap = matmul(T, b%xAxisAziFlap)
        b%yAxisAziFlap = matmul(T, b%yAxisAziFlap)
        b%zAxisAziFlap = matmul(T, b%zAxisAziFlap)
      end associate
    enddo
    r%shaftAxis = matmul(T, r%shaftAxis)
    r%xAxisBody = matmul(T, r%xAxisBody)
    r%yAxisBody = matmul(T, r%yAxisBody)
    r%zAxisBody = matmul(T, r%zAxisBody)
    r%hubCoords = matmul(T, r%hubCoords - org) + org
    r%cgCoords = matmul(T, r%cgCoords - org) + org
  end subroutine rotor_rot_pts

  function rotor_gettheta(r, psi, ib) result(th)
    !! Collective plus cyclic pitch, classdef.f90:4114-4136
    type(rotor_class), intent(in) :: r
    real(dp), intent(in) :: psi
    integer, intent(in) :: ib
    real(dp) :: th, off
    off = twoPi/r%nb*(ib - 1)
    th = r%controlPitch(1) + r%controlPitch(2)*cos(psi + off) + &
      & r%controlPitch(3)*sin(psi + off)
  end function rotor_gettheta

  subroutine rotor_rot_advance(r, dpsi, nopitch)
    !! Azimuthal advance with pitch schedule, classdef.f90:4254-4291
    type(rotor_class), intent(inout) :: r
    real(dp), intent(in) :: dpsi
    logical, intent(in) :: nopitch
    integer :: ib
    real(dp) :: thNext
    r%psi = r%psi + dpsi
    do ib = 1, r%nb
      call blade_rotate(r%blade(ib), dpsi, r%shaftAxis, r%hubCoords, 1)
      r%blade(ib)%psi = r%blade(ib)%psi + dpsi
      if (.not. nopitch) then
        thNext = rotor_gettheta(r, r%psi, ib)
        call blade_rot_pitch(r%blade(ib), thNext - r%blade(ib)%theta)
        r%blade(ib)%theta = thNext
      endif
    enddo
  end subroutine rotor_rot_advance

  subroutine rotor_move(r, d)
    type(rotor_class), intent(inout) :: r
    real(dp), intent(in) :: d(3)
    integer :: ib
    do ib = 1, r%nb
      call blade_move(r%blade(ib), d)
    enddo
    r%hubCoords = r%hubCoords + d
    r%cgCoords = r%cgCoords + d
  end subroutine rotor_move

  ! ------------------------------------------------------------------
  ! Geometry initialisation, classdef.f90:2769-3891
  ! ------------------------------------------------------------------
  subroutine image_params(r, src, nt)
    !! The parameters an image rotor takes from its (initialised) source,
    !! mirrored in the plane where coordinate imagePlane is zero
    !! (rotor_init, classdef.f90:2798-2955).  As there, only the hub and cg
    !! are mirrored for every plane; the rotation sense, start azimuth,
    !! pitch and body motion follow the per-plane rules of 2922-2933, and
    !! the body axes, mirrored there, are reset to x, y, z later (3034-3036),
    !! which leaves them as they are here.  prescWakeAfterTruncNt is not
    !! copied (the image never builds a prescribed wake: it does not convect).
    type(rotor_class), intent(inout) :: r
    type(rotor_class), intent(in) :: src
    integer, intent(in) :: nt
    integer :: p
    p = r%imagePlane
    r%nb = src%nb; r%propConvention = src%propConvention
    r%spanSpacing = src%spanSpacing; r%chordSpacing = src%chordSpacing
    r%nCamberFiles = src%nCamberFiles
    r%camberSectionLimit = src%camberSectionLimit; r%camberFile = src%camberFile
    r%nc = src%nc; r%ns = src%ns; r%nNwake = src%nNwake
    r%hubCoords = src%hubCoords; r%cgCoords = src%cgCoords; r%fromCoords = src%fromCoords
    r%hubCoords(p) = -1._dp*r%hubCoords(p)
    r%cgCoords(p) = -1._dp*r%cgCoords(p)
    r%pts = src%pts; r%radius = src%radius; r%root_cut = src%root_cut; r%chord = src%chord
    r%preconeAngle = src%preconeAngle
    r%Omega = src%Omega; r%shaftAxis = src%shaftAxis
    r%controlPitch = src%controlPitch; r%thetaTwist = src%thetaTwist
    r%customTrajectorySwitch = src%customTrajectorySwitch
    r%axisymmetrySwitch = src%axisymmetrySwitch
    r%velBody = src%velBody; r%omegaBody = src%omegaBody
    r%pivotLE = src%pivotLE; r%flapHinge = src%flapHinge
    r%spanwiseLiftSwitch = src%spanwiseLiftSwitch; r%symmetricTau = src%symmetricTau
    r%apparentViscCoeff = src%apparentViscCoeff; r%decayCoeff = src%decayCoeff
    r%wakeTruncateNt = src%wakeTruncateNt
    r%prescWakeNt = src%prescWakeNt; r%prescWakeGenNt = src%prescWakeGenNt
    r%spanwiseCore = src%spanwiseCore
    if (allocated(r%streamwiseCoreVec)) deallocate (r%streamwiseCoreVec)
    allocate (r%streamwiseCoreVec(r%ns + 1))
    r%streamwiseCoreVec = src%streamwiseCoreVec
    r%rollupStartRadius = src%rollupStartRadius; r%rollupEndRadius = src%rollupEndRadius
    r%rollupStart = src%rollupStart; r%rollupEnd = src%rollupEnd
    r%initWakeVel = src%initWakeVel; r%psiStart = src%psiStart; r%skewLimit = src%skewLimit
    r%flapInitial = src%flapInitial; r%dflapInitial = src%dflapInitial; r%Iflap = src%Iflap
    r%cflap = src%cflap; r%kflap = src%kflap; r%MflapConstant = src%MflapConstant
    r%dpitch = src%dpitch
    r%inflowPlotSwitch = src%inflowPlotSwitch; r%gammaPlotSwitch = src%gammaPlotSwitch
    r%skewPlotSwitch = src%skewPlotSwitch
    r%forceCalcSwitch = src%forceCalcSwitch
    select case (p)
    case (1)
      r%Omega = -1._dp*r%Omega
    case (2)
      r%Omega = -1._dp*r%Omega
      r%psiStart = r%psiStart + pi
    case (3)
      r%controlPitch = -1._dp*r%controlPitch
      r%velBody(3) = -1._dp*r%velBody(3)
      r%omegaBody(1) = -1._dp*r%omegaBody(1)
      r%omegaBody(2) = -1._dp*r%omegaBody(2)
    end select
    if (r%customTrajectorySwitch == 1) then
      allocate (r%velBodyHistory(3, nt), r%omegaBodyHistory(3, nt))
      r%velBodyHistory = src%velBodyHistory
      r%omegaBodyHistory = src%omegaBodyHistory
      r%velBodyHistory(p, :) = -1._dp*r%velBodyHistory(p, :)
      if (p == 3) then
        r%omegaBodyHistory(1, :) = -1._dp*r%omegaBodyHistory(1, :)
        r%omegaBodyHistory(2, :) = -1._dp*r%omegaBodyHistory(2, :)
      else
        r%omegaBodyHistory(3, :) = -1._dp*r%omegaBodyHistory(3, :)
      endif
    endif
  end subroutine image_params

  subroutine rotor_init(r, ir, density, dt, nt, sw, dir, src)
    type(rotor_class), intent(inout) :: r
    integer, intent(in) :: ir
    real(dp), intent(in) :: density
    real(dp), intent(inout) :: dt
    integer, intent(inout) :: nt
    type(switches_class), intent(inout) :: sw
    character(len=*), intent(in) :: dir
    type(rotor_class), intent(in), optional :: src  !! the source of an image rotor
    real(dp), allocatable :: xVec(:), yVec(:), grid(:, :, :)
    real(dp) :: dxdymin, velShed, xs(4), dummy
    integer :: ib, i, j, nx, ny, nz, u, ic, is, kb

    write (r%id, '(I0.2)') ir
    if (r%surfaceType < 0) then
      if (.not. present(src)) call host_fail('ERROR: image rotor without its source rotor')
      call image_params(r, src, nt)
    endif
    allocate (xVec(r%nc + 1), yVec(r%ns + 1), grid(3, r%nc + 1, r%ns + 1))
    if (dt < 0._dp) then
      if (abs(r%Omega) < eps) then
        dt = abs(dt)*r%chord/norm2(r%velBody)
      else
        dt = twoPi*abs(dt)/abs(r%Omega)
      endif
    endif
    if (abs(dt) <= eps) then
      if (abs(r%Omega) < eps) then
        dt = (r%chord/r%nc)/norm2(r%velBody)
      else
        dt = 5._dp*degToRad/abs(r%Omega)
      endif
    endif
    if (nt <= 0) then
      if (nt == 0) nt = -10
      call to_chords_revs(r, nt, dt)
    endif
    if (sw%slowStart /= 0) call to_chords_revs(r, sw%slowStartNt, dt)
    ! plot intervals in chords or revolutions become steps (classdef.f90:3001-3003);
    ! they are reported in rXXParams.json, no wake or grid plots are written
    call to_chords_revs(r, sw%wakeTipPlot, dt)
    call to_chords_revs(r, sw%wakePlot, dt)
    call to_chords_revs(r, sw%gridPlot, dt)
    call to_chords_revs(r, r%wakeTruncateNt, dt)
    call to_chords_revs(r, r%prescWakeAfterTruncNt, dt)
    call to_chords_revs(r, r%prescWakeGenNt, dt)
    call to_chords_revs(r, r%nNwake, dt)
    call to_chords_revs(r, r%inflowPlotSwitch, dt)
    call to_chords_revs(r, r%gammaPlotSwitch, dt)
    call to_chords_revs(r, r%skewPlotSwitch, dt)
    ! the prescribed far wake starts prescWakeAfterTruncNt steps after the
    ! truncation, counted before wakeTruncateNt is raised to nNwake + 1
    ! (classdef.f90:3013-3021)
    if (r%wakeTruncateNt > 0 .and. r%prescWakeAfterTruncNt > 0) then
      r%prescWakeNt = r%wakeTruncateNt + r%prescWakeAfterTruncNt
    else
      r%prescWakeNt = 0
    endif
    if (r%wakeTruncateNt > 0) r%wakeTruncateNt = max(r%wakeTruncateNt, r%nNwake + 1)
    if (r%surfaceType == 0) r%surfaceType = 1
    if (r%nNwake > 0 .and. r%nNwake < 2) call host_fail('ERROR: Atleast 2 near wake rows mandatory')
    r%nNwake = min(r%nNwake, nt)
    if (r%wakeTruncateNt == 0) then
      r%nFwake = nt - r%nNwake
    else
      r%nFwake = r%wakeTruncateNt - r%nNwake
    endif
    allocate (r%ageAzN(max(r%nNwake, 1)), r%ageAzF(max(r%nFwake, 1)))
    r%ageAzN = 0._dp; r%ageAzF = 0._dp
    ! custom trajectory: body velocity and rate of every step from
    ! trajectoryXX.in, nt lines of six values (classdef.f90:3064-3080)
    if (r%customTrajectorySwitch == 1 .and. r%surfaceType >= 0) then
      block
        integer :: u, i
        allocate (r%velBodyHistory(3, nt), r%omegaBodyHistory(3, nt))
        call open_old(u, trim(dir)//'/trajectory'//r%id//'.in')
        do i = 1, nt
          read (u, *) r%velBodyHistory(:, i), r%omegaBodyHistory(:, i)
        enddo
        close (u)
      end block
    endif
    r%nbConvect = r%nb
    if (r%axisymmetrySwitch == 1) r%nbConvect = 1

    allocate (r%blade(r%nb))
    allocate (r%gamVec(r%nc*r%ns*r%nb), r%gamVecPrev(r%nc*r%ns*r%nb), r%RHS(r%nc*r%ns*r%nb))
    r%gamVec = 0._dp; r%gamVecPrev = 0._dp; r%RHS = 0._dp
    do ib = 1, r%nb
      associate (b => r%blade(ib))
        allocate (b%wiP(r%nc, r%ns))
        allocate (b%secCP(3, r%ns), b%secTauCapChord(3, r%ns), b%secTauCapSpan(3, r%ns))
        allocate (b%secNormalVec(3, r%ns), b%secChordwiseResVel(3, r%ns))
        allocate (b%secLift(3, r%ns), b%secLiftDir(3, r%ns), b%secDragDir(3, r%ns))
        allocate (b%secForceInertial(3, r%ns), b%secLiftUnsteady(3, r%ns))
        allocate (b%secLiftInPlane(3, r%ns), b%secLiftOutPlane(3, r%ns))
        allocate (b%secChord(r%ns), b%secArea(r%ns), b%secCL(r%ns), b%secCLu(r%ns))
        allocate (b%secAlpha(r%ns), b%secMflapArm(r%ns))
        allocate (b%secPhi(r%ns), b%secTheta(r%ns), b%secViz(r%ns), b%secVix(r%ns))
        b%secCP = 0._dp; b%secTauCapChord = 0._dp; b%secTauCapSpan = 0._dp
        b%secNormalVec = 0._dp; b%secChordwiseResVel = 0._dp; b%secLift = 0._dp
        b%secLiftDir = 0._dp; b%secDragDir = 0._dp; b%secForceInertial = 0._dp
        b%secLiftUnsteady = 0._dp; b%secLiftInPlane = 0._dp; b%secLiftOutPlane = 0._dp
        b%secChord = 0._dp; b%secArea = 0._dp; b%secCL = 0._dp; b%secCLu = 0._dp
        b%secAlpha = 0._dp; b%secMflapArm = 0._dp
        b%secPhi = 0._dp; b%secTheta = 0._dp; b%secViz = 0._dp; b%secVix = 0._dp
      end associate
    enddo

    ! an image's copied parameters are converted already (classdef.f90:3126-3137)
    if (r%surfaceType >= 0) then
      r%controlPitch = r%controlPitch*degToRad
      r%pts = r%pts*degToRad
      r%preconeAngle = r%preconeAngle*degToRad
      r%psiStart = r%psiStart*degToRad
      r%thetaTwist = r%thetaTwist*degToRad
      r%spanwiseCore = r%spanwiseCore*r%chord
      r%streamwiseCoreVec = r%streamwiseCoreVec*r%chord
      r%rollupStart = ceiling(r%rollupStartRadius*r%ns)
      r%rollupEnd = floor(r%rollupEndRadius*r%ns)
    endif

    ! Panel corners
    if (r%geometryFile(1:1) == '0') then
      if (r%Omega >= 0._dp) then
        xVec = spaced(r%chordSpacing, -r%chord, 0._dp, r%nc + 1)
      else
        xVec = spaced(r%chordSpacing, r%chord, 0._dp, r%nc + 1)
      endif
      yVec = spaced(r%spanSpacing, r%root_cut*r%radius, r%radius, r%ns + 1)
      do j = 1, r%ns + 1
        do i = 1, r%nc + 1
          grid(:, i, j) = [xVec(i), yVec(j), 0._dp]
        enddo
      enddo
      if (r%nCamberFiles > 0) grid(3, :, :) = wing_camber(r, dir, xVec, yVec)
    else
      call open_old(u, trim(dir)//'/'//trim(r%geometryFile))
      read (u, *) nx, ny, nz
      if (nz > 1 .or. nx /= r%nc + 1 .or. ny /= r%ns + 1) &
        call host_fail('ERROR: Wrong or conflicting data in PLOT3D file')
      read (u, *) ((grid(1, i, j), i=1, nx), j=1, ny), ((grid(2, i, j), i=1, nx), j=1, ny), &
        & ((grid(3, i, j), i=1, nx), j=1, ny)
      close (u)
    endif
    ! an image in the plane z = 0 starts from the blade mirrored in z
    ! (classdef.f90:3228-3230, 3990-3997)
    if (r%surfaceType < 0 .and. r%imagePlane == 3) grid(3, :, :) = -1._dp*grid(3, :, :)
    do ib = 1, r%nb
      do j = 1, r%ns
        do i = 1, r%nc
          r%blade(ib)%wiP(i, j)%PC(:, 1) = grid(:, i, j)
          r%blade(ib)%wiP(i, j)%PC(:, 2) = grid(:, i + 1, j)
          r%blade(ib)%wiP(i, j)%PC(:, 3) = grid(:, i + 1, j + 1)
          r%blade(ib)%wiP(i, j)%PC(:, 4) = grid(:, i, j + 1)
        enddo
      enddo
    enddo

    do ib = 1, r%nb
      associate (b => r%blade(ib))
        b%xAxis = xAxis; b%yAxis = yAxis; b%zAxis = zAxis
        b%xAxisAzi = xAxis; b%yAxisAzi = yAxis; b%zAxisAzi = zAxis
        b%xAxisAziFlap = xAxis; b%yAxisAziFlap = yAxis; b%zAxisAziFlap = zAxis
        b%flapOrigin = b%yAxis*r%radius*r%flapHinge
        do j = 1, r%ns
          b%secTauCapChord(:, j) = unit3((b%wiP(r%nc, j)%PC(:, 3) + b%wiP(r%nc, j)%PC(:, 2) - &
            & (b%wiP(1, j)%PC(:, 4) + b%wiP(1, j)%PC(:, 1)))*0.5_dp)
          b%secTauCapSpan(:, j) = yAxis
          b%secNormalVec(:, j) = sign(1._dp, r%Omega)*unit3(cross3( &
            & b%wiP(r%nc, j)%PC(:, 2) - b%wiP(1, j)%PC(:, 4), &
            & b%wiP(r%nc, j)%PC(:, 3) - b%wiP(1, j)%PC(:, 1)))
        enddo
        ! vortex ring corners a quarter panel downstream of the panel corners
        do j = 1, r%ns
          do i = 1, r%nc
            associate (w => b%wiP(i, j))
              xs(1) = (w%PC(1, 2) - w%PC(1, 1))*0.25_dp
              xs(4) = (w%PC(1, 3) - w%PC(1, 4))*0.25_dp
              if (i < r%nc) then
                xs(2) = (b%wiP(i + 1, j)%PC(1, 2) - w%PC(1, 2))*0.25_dp
                xs(3) = (b%wiP(i + 1, j)%PC(1, 3) - w%PC(1, 3))*0.25_dp
              else
                xs(2) = 0._dp
                xs(3) = 0._dp
              endif
              do is = 1, 4
                w%vr%fc(:, is) = [w%PC(1, is) + xs(is), w%PC(2, is), w%PC(3, is)]
              enddo
            end associate
          enddo
        enddo
        dxdymin = huge(1._dp)
        do is = 1, r%ns
          do ic = 1, r%nc
            dxdymin = min(dxdymin, abs(norm2(b%wiP(ic, is)%PC(:, 2) - b%wiP(ic, is)%PC(:, 1))), &
              & abs(norm2(b%wiP(ic, is)%PC(:, 3) - b%wiP(ic, is)%PC(:, 2))))
          enddo
        enddo
        ! shed the trailing vortex row slightly downstream (single-precision
        ! 0.05 factor, as in the reference's literal)
        if (abs(r%Omega) > eps) then
          velShed = min(0.05*abs(r%Omega)*norm2(b%wiP(r%nc, r%ns)%vr%fc(:, 2) - r%hubCoords), &
            & 0.125_dp*r%chord/dt)
        else
          velShed = 0.3_dp*norm2(-1._dp*r%velBody)
        endif
        do j = 1, r%ns
          b%wiP(r%nc, j)%vr%fc(:, 2) = b%wiP(r%nc, j)%vr%fc(:, 2) + &
            & [sign(1._dp, r%Omega)*velShed*dt, 0._dp, 0._dp]
          b%wiP(r%nc, j)%vr%fc(:, 3) = b%wiP(r%nc, j)%vr%fc(:, 3) + &
            & [sign(1._dp, r%Omega)*velShed*dt, 0._dp, 0._dp]
        enddo
        do j = 1, r%ns
          do i = 1, r%nc
            associate (w => b%wiP(i, j))
              w%CP = ((w%PC(:, 1) + w%PC(:, 4))*0.25_dp + (w%PC(:, 2) + w%PC(:, 3))*0.75_dp)*0.5_dp
              w%nCap = unit3(cross3(w%PC(:, 3) - w%PC(:, 1), w%PC(:, 4) - w%PC(:, 2)))
              if (sign(1._dp, r%Omega) < 0._dp) w%nCap = -1._dp*w%nCap
              w%tauCapChord = unit3(0.5_dp*((w%PC(:, 2) + w%PC(:, 3)) - (w%PC(:, 1) + w%PC(:, 4))))
              w%tauCapSpan = unit3(0.5_dp*((w%PC(:, 3) + w%PC(:, 4)) - (w%PC(:, 2) + w%PC(:, 1))))
              w%panelArea = 0.5_dp*norm2(cross3(w%PC(:, 3) - w%PC(:, 1), w%PC(:, 4) - w%PC(:, 2)))
              w%meanSpan = 0.5_dp*(norm2(w%PC(:, 4) - w%PC(:, 1)) + norm2(w%PC(:, 3) - w%PC(:, 2)))
              w%meanChord = 0.5_dp*(norm2(w%PC(:, 2) - w%PC(:, 1)) + norm2(w%PC(:, 3) - w%PC(:, 4)))
            end associate
          enddo
        enddo
        do is = 1, r%ns
          b%secArea(is) = sum(b%wiP(:, is)%panelArea)
          b%secChord(is) = norm2(0.5_dp*((b%wiP(1, is)%PC(:, 1) + b%wiP(1, is)%PC(:, 4)) - &
            & (b%wiP(r%nc, is)%PC(:, 2) + b%wiP(r%nc, is)%PC(:, 3))))
        enddo
        do j = 1, r%ns
          b%secTauCapSpan(:, j) = b%yAxis
          do i = 1, r%nc
            b%wiP(i, j)%tauCapSpan = b%yAxis
          enddo
        enddo
        if (r%symmetricTau == 1) then
          do j = 1, r%ns/2
            b%secTauCapSpan(:, j) = -1._dp*b%secTauCapSpan(:, j)
            do i = 1, r%nc
              b%wiP(i, j)%tauCapSpan = -1._dp*b%wiP(i, j)%tauCapSpan
            enddo
          enddo
        endif
        call blade_sec_locations(b, 0.5_dp, r%flapHinge*r%radius)
        b%pivotLE = r%pivotLE
        do j = 1, r%ns
          do i = 1, r%nc
            b%wiP(i, j)%vr%gam = 0._dp
            b%wiP(i, j)%vr%rVc = min(r%spanwiseCore, dxdymin*0.1_dp)
          enddo
          b%wiP(r%nc, j)%vr%rVc(2) = r%spanwiseCore
        enddo
      end associate
    enddo

    if (r%surfaceType < 0) then
      ! the image's normals are its source's mirrored in the image plane,
      ! taken before the image's own hub, precone and azimuth rotations
      ! (classdef.f90:3434-3445); axisymmetric blades copy blade 1's
      ! (3566-3572)
      do ib = 1, r%nb
        kb = ib
        if (r%axisymmetrySwitch == 1) kb = 1
        do j = 1, r%ns
          do i = 1, r%nc
            r%blade(ib)%wiP(i, j)%nCap = src%blade(kb)%wiP(i, j)%nCap
            r%blade(ib)%wiP(i, j)%nCap(r%imagePlane) = &
              & -1._dp*src%blade(kb)%wiP(i, j)%nCap(r%imagePlane)
          enddo
        enddo
      enddo
    endif
    do ib = 1, r%nb
      call blade_move(r%blade(ib), r%hubCoords - r%fromCoords)
    enddo
    if (abs(r%preconeAngle) > eps) then
      do ib = 1, r%nb
        call blade_rotate(r%blade(ib), r%preconeAngle, r%blade(ib)%xAxisAzi, &
          & r%blade(ib)%flapOrigin, 2)
      enddo
    endif
    do ib = 2, r%nb
      call blade_rotate(r%blade(ib), sign(1._dp, r%Omega)*twoPi/r%nb*(ib - 1), r%shaftAxis, &
        & r%hubCoords, 1)
    enddo
    call rotor_rot_pts(r, r%pts, r%cgCoords)
    call rotor_rot_advance(r, sign(1._dp, r%Omega)*r%psiStart, .true.)

    if (abs(r%Omega) > eps) then
      if (r%propConvention == 0) then
        r%nonDimforceDenominator = density*(pi*r%radius**2)*(r%radius*r%Omega)**2
      else
        r%nonDimforceDenominator = density*(r%Omega/twoPi)**2*(2._dp*r%radius)**4
      endif
    else
      r%nonDimforceDenominator = 0.5_dp*density*(r%radius*(1._dp - r%root_cut)*r%chord)* &
        & dot_product(r%velBody, r%velBody)
    endif
    dummy = 0._dp
  end subroutine rotor_init

  subroutine blade_sec_locations(b, frac, hingeR)
    !! Section reference points and flap arms, classdef.f90:2267-2304
    type(blade_class), intent(inout) :: b
    real(dp), intent(in) :: frac, hingeR
    real(dp) :: xz(2, size(b%wiP, 1) + 1), vPC(3), vLE(3), xcp, zcp
    integer :: is, ic, nc
    nc = size(b%wiP, 1)
    do is = 1, size(b%wiP, 2)
      vLE = 0.5_dp*(b%wiP(1, is)%PC(:, 1) + b%wiP(1, is)%PC(:, 4))
      xz(:, 1) = 0._dp
      do ic = 1, nc
        vPC = 0.5_dp*(b%wiP(ic, is)%PC(:, 2) + b%wiP(ic, is)%PC(:, 3)) - vLE
        xz(1, ic + 1) = norm2(proj3(vPC, b%secTauCapChord(:, is)))
        xz(2, ic + 1) = dot_product(vPC, b%secNormalVec(:, is))
      enddo
      xcp = norm2(vPC)*frac
      zcp = pwl(xz(1, :), xz(2, :), xcp)
      b%secCP(:, is) = vLE + xcp*b%secTauCapChord(:, is) + zcp*b%secNormalVec(:, is)
      b%secMflapArm(is) = norm2(proj3(b%secCP(:, is), b%yAxis)) - hingeR
    enddo
  end subroutine blade_sec_locations

  ! ------------------------------------------------------------------
  ! Device hand-off of the wing
  ! ------------------------------------------------------------------
  subroutine rotor_upload_wing(ctx, r)
    type(c_ptr), intent(in) :: ctx
    type(rotor_class), intent(in) :: r
    real(c_double) :: wc(3, 4, r%nc, r%ns, r%nb), wr(4, r%nc, r%ns, r%nb)
    real(c_double) :: wg(r%nc, r%ns, r%nb)
    integer :: ib, i, j
    do ib = 1, r%nb
      do j = 1, r%ns
        do i = 1, r%nc
          wc(:, :, i, j, ib) = r%blade(ib)%wiP(i, j)%vr%fc
          wr(:, i, j, ib) = r%blade(ib)%wiP(i, j)%vr%rVc
          wg(i, j, ib) = r%blade(ib)%wiP(i, j)%vr%gam
        enddo
      enddo
    enddo
    call vlm_check(ctx, vlm_rotor_set_wing(ctx, r%dev, wc, wr, wg), 'rotor_upload_wing')
    call vlm_check(ctx, vlm_rotor_set_axis(ctx, r%dev, r%shaftAxis, r%hubCoords), &
      & 'rotor_set_axis')
    call rotor_upload_panels(ctx, r)
  end subroutine rotor_upload_wing

  subroutine rotor_upload_panels(ctx, r)
    !! The panel, section and blade geometry of the right-hand side and the
    !! loads on the device (vlm_rotor_set_panels records, ring order)
    type(c_ptr), intent(in) :: ctx
    type(rotor_class), intent(in) :: r
    real(c_double) :: pn(15, r%nc, r%ns, r%nb), sc(16, r%ns, r%nb), bl(9, r%nb)
    integer :: ib, i, j
    do ib = 1, r%nb
      associate (b => r%blade(ib))
        do j = 1, r%ns
          do i = 1, r%nc
            associate (w => b%wiP(i, j))
              pn(1:3, i, j, ib) = w%CP
              pn(4:6, i, j, ib) = w%nCap
              pn(7:9, i, j, ib) = w%tauCapChord
              pn(10:12, i, j, ib) = w%tauCapSpan
              pn(13, i, j, ib) = w%panelArea
              pn(14, i, j, ib) = w%meanChord
              pn(15, i, j, ib) = w%meanSpan
            end associate
          enddo
          sc(1:3, j, ib) = b%secCP(:, j)
          sc(4:6, j, ib) = b%wiP(1, j)%PC(:, 1)
          sc(7:9, j, ib) = b%secTauCapChord(:, j)
          sc(10:12, j, ib) = b%secTauCapSpan(:, j)
          sc(13:15, j, ib) = b%secNormalVec(:, j)
          sc(16, j, ib) = b%secArea(j)
        enddo
        bl(1:3, ib) = b%xAxisAziFlap
        bl(4:6, ib) = b%yAxisAziFlap
        bl(7:9, ib) = b%zAxisAziFlap
      end associate
    enddo
    call vlm_check(ctx, vlm_rotor_set_panels(ctx, r%dev, pn, sc, bl), 'set_panels')
  end subroutine rotor_upload_panels

  subroutine rotor_cp_arrays(r, cp, ncap)
    type(rotor_class), intent(in) :: r
    real(c_double), intent(out) :: cp(3, r%nc, r%ns, r%nb), ncap(3, r%nc, r%ns, r%nb)
    integer :: ib, i, j
    do ib = 1, r%nb
      do j = 1, r%ns
        do i = 1, r%nc
          cp(:, i, j, ib) = r%blade(ib)%wiP(i, j)%CP
          ncap(:, i, j, ib) = r%blade(ib)%wiP(i, j)%nCap
        enddo
      enddo
    enddo
  end subroutine rotor_cp_arrays

  subroutine rotor_map_gam(r)
    !! classdef.f90:4181-4196
    type(rotor_class), intent(inout) :: r
    integer :: ib, i, j
    do ib = 1, r%nbConvect
      do j = 1, r%ns
        do i = 1, r%nc
          r%blade(ib)%wiP(i, j)%vr%gam = r%gamVec(i + r%nc*(j - 1) + r%ns*r%nc*(ib - 1))
        enddo
      enddo
    enddo
    if (r%axisymmetrySwitch == 1) then
      do ib = 2, r%nb
        r%blade(ib)%wiP%vr%gam = r%blade(1)%wiP%vr%gam
      enddo
    endif
  end subroutine rotor_map_gam

  ! ------------------------------------------------------------------
  ! RHS, solve and loads
  ! ------------------------------------------------------------------
  subroutine compute_rhs_and_solve(ctx, rotor, nr, with_wake)
    !! main.f90:124-202 (initial) and 521-603 (every step), enqueued on the
    !! device: per rotor the right-hand side from the rigid motion and the
    !! wake and (other) wing velocities at the collocation points, summed
    !! in the reference's order (main.f90:549-561), then the solve of that
    !! right-hand side; an image mirrors its source's collocation velocities
    !! and circulation instead (main.f90:181-193, 595-599).  gamVec and RHS
    !! arrive at vlm_wait, after which finish_solve maps gamVec onto the
    !! panels (the device wing circulation is already set).
    type(c_ptr), intent(in) :: ctx
    type(rotor_class), intent(inout), target :: rotor(:)
    integer, intent(in) :: nr
    logical, intent(in) :: with_wake
    integer :: ir, jr, nt
    integer(c_int) :: tir(2*nr), tset(2*nr), tfl(2*nr)
    real(c_double) :: kin(10)

    do ir = 1, nr
      associate (r => rotor(ir))
        if (r%surfaceType < 0) then
          call vlm_check(ctx, vlm_rotor_mirror_velcp(ctx, r%dev), 'mirror_velcp')
          r%RHS = 0._dp
          cycle
        endif
        nt = 0
        do jr = 1, nr
          if (with_wake .and. rotor(jr)%nNwake > 0) then
            nt = nt + 1; tir(nt) = rotor(jr)%dev; tset(nt) = VLM_SET_WAKE; tfl(nt) = 0
          endif
          if (jr /= ir) then
            nt = nt + 1; tir(nt) = rotor(jr)%dev; tset(nt) = VLM_SET_WING; tfl(nt) = 0
          endif
        enddo
        kin(1:3) = r%velBody
        kin(4:6) = r%omegaBody
        kin(7:9) = r%cgCoords
        kin(10) = r%omegaSlow
        call vlm_check(ctx, vlm_rotor_rhs(ctx, r%dev, int(nt, c_int), tir, tset, tfl, kin, &
          & c_loc(r%RHS)), 'RHS')
      end associate
    enddo
    do ir = 1, nr
      rotor(ir)%gamVecPrev = rotor(ir)%gamVec
      if (rotor(ir)%surfaceType < 0) then
        ! rotor_mirrorGamma: the device wing circulation now, gamVec in
        ! finish_solve (main.f90:595-599)
        call vlm_check(ctx, vlm_rotor_mirror_gamma(ctx, rotor(ir)%dev), 'mirror_gamma')
        cycle
      endif
      call vlm_check(ctx, vlm_rotor_solve_deferred(ctx, rotor(ir)%dev, c_null_ptr, &
        & c_loc(rotor(ir)%gamVec), 3_c_int, 1.e-15_c_double, c_loc(rotor(ir)%solve_iters)), &
        & 'rotor_solve')
    enddo
  end subroutine compute_rhs_and_solve

  subroutine finish_solve(rotor, nr)
    !! After vlm_wait: rotor_map_gam on the host panels
    type(rotor_class), intent(inout) :: rotor(:)
    integer, intent(in) :: nr
    integer :: ir
    do ir = 1, nr
      ! an image's gamVec is its source's negated (rotor_mirrorGamma)
      if (rotor(ir)%surfaceType < 0) rotor(ir)%gamVec = -1._dp*rotor(rotor(ir)%imageRotorNum)%gamVec
      call rotor_map_gam(rotor(ir))
    enddo
  end subroutine finish_solve

  subroutine restart_io(ctx, u, rotor, nr, t, writing)
    !! One restart record set (the reference writes t and its whole rotor
    !! array, main.f90:1442-1450, and reads them back, 390-398): here t,
    !! every host field a step changes (rigid-motion frames, blade panels
    !! and sections, circulation, loads) and the device state snapshot of
    !! each rotor (vlm_rotor_get_state / vlm_rotor_set_state).
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: u, nr
    type(rotor_class), intent(inout) :: rotor(:)
    real(dp), intent(inout) :: t
    logical, intent(in) :: writing
    integer(c_int8_t), allocatable :: blob(:)
    integer(c_int64_t) :: nbytes, nfile
    integer :: ir, ib
    if (writing) then
      write (u) t
    else
      read (u) t
    endif
    do ir = 1, nr
      associate (r => rotor(ir))
        if (writing) then
          write (u) r%rowNear, r%rowFar, r%omegaSlow, r%psi, r%shaftAxis, r%hubCoords, &
            & r%cgCoords, r%fromCoords, r%pts, r%xAxisBody, r%yAxisBody, r%zAxisBody, &
            & r%velBody, r%omegaBody, r%controlPitch, r%gamVec, r%gamVecPrev, r%RHS, r%lift, &
            & r%drag, r%forceInertial, r%liftUnsteady, r%ageAzN, r%ageAzF
        else
          read (u) r%rowNear, r%rowFar, r%omegaSlow, r%psi, r%shaftAxis, r%hubCoords, &
            & r%cgCoords, r%fromCoords, r%pts, r%xAxisBody, r%yAxisBody, r%zAxisBody, &
            & r%velBody, r%omegaBody, r%controlPitch, r%gamVec, r%gamVecPrev, r%RHS, r%lift, &
            & r%drag, r%forceInertial, r%liftUnsteady, r%ageAzN, r%ageAzF
        endif
        do ib = 1, r%nb
          associate (b => r%blade(ib))
            if (writing) then
              write (u) b%wiP, b%xAxis, b%yAxis, b%zAxis, b%xAxisAzi, b%yAxisAzi, b%zAxisAzi, &
                & b%xAxisAziFlap, b%yAxisAziFlap, b%zAxisAziFlap, b%theta, b%psi, b%pivotLE, &
                & b%flapOrigin, b%lift, b%drag, b%forceInertial, b%liftUnsteady
              write (u) b%secCP, b%secTauCapChord, b%secTauCapSpan, b%secNormalVec, &
                & b%secChordwiseResVel, b%secLift, b%secLiftDir, b%secDragDir, &
                & b%secForceInertial, b%secLiftUnsteady, b%secLiftInPlane, b%secLiftOutPlane, &
                & b%secChord, b%secArea, b%secCL, b%secCLu, b%secAlpha, b%secMflapArm, &
                & b%secPhi, b%secTheta, b%secViz, b%secVix
            else
              read (u) b%wiP, b%xAxis, b%yAxis, b%zAxis, b%xAxisAzi, b%yAxisAzi, b%zAxisAzi, &
                & b%xAxisAziFlap, b%yAxisAziFlap, b%zAxisAziFlap, b%theta, b%psi, b%pivotLE, &
                & b%flapOrigin, b%lift, b%drag, b%forceInertial, b%liftUnsteady
              read (u) b%secCP, b%secTauCapChord, b%secTauCapSpan, b%secNormalVec, &
                & b%secChordwiseResVel, b%secLift, b%secLiftDir, b%secDragDir, &
                & b%secForceInertial, b%secLiftUnsteady, b%secLiftInPlane, b%secLiftOutPlane, &
                & b%secChord, b%secArea, b%secCL, b%secCLu, b%secAlpha, b%secMflapArm, &
                & b%secPhi, b%secTheta, b%secViz, b%secVix
            endif
          end associate
        enddo
        call vlm_check(ctx, vlm_rotor_state_bytes(ctx, r%dev, nbytes), 'state_bytes')
        allocate (blob(nbytes))
        if (writing) then
          call vlm_check(ctx, vlm_rotor_get_state(ctx, r%dev, blob, nbytes), 'get_state')
          write (u) nbytes
          write (u) blob
        else
          read (u) nfile
          if (nfile /= nbytes) call host_fail('ERROR: restart file does not match this case')
          read (u) blob
          call vlm_check(ctx, vlm_rotor_set_state(ctx, r%dev, blob, nbytes), 'set_state')
        endif
        deallocate (blob)
      end associate
    enddo
  end subroutine restart_io

  subroutine skew_file(ctx, dir, timestamp, r)
    !! calc_skew and skew2file (classdef.f90:704-747 and 4919-4937,
    !! libPostprocess.f90:786-812): the skew of every active near-wake ring,
    !! |x1 . x2| / sqrt(|x1|^2 |x2|^2) of its bimedians x1 = p3+p4-p1-p2 and
    !! x2 = p4+p1-p2-p3 (0 where |Γ| <= eps), on the convected blades and
    !! copied to the others of an axisymmetric rotor; per blade, the maximum
    !! and the mean over the span of each row from the oldest to rowNear
    type(c_ptr), intent(in) :: ctx
    character(len=*), intent(in) :: dir, timestamp
    type(rotor_class), intent(in) :: r
    real(c_double), allocatable, target :: nodes(:, :, :, :), le(:, :, :), gam(:, :, :), &
      & rc4(:, :, :, :), ffc(:, :, :, :), fgam(:, :), frc(:, :)
    real(dp) :: skew(r%nNwake, r%ns, r%nb), p(3, 4), x1(3), x2(3)
    character(len=2) :: bid
    integer :: ib, irow, icol, u, nF
    nF = max(r%nFwake, 1)
    allocate (nodes(3, r%ns + 1, r%nNwake, r%nb), le(3, r%ns + 1, r%nb), &
      & gam(r%ns, r%nNwake, r%nb), rc4(4, r%ns, r%nNwake, r%nb), ffc(3, 2, nF, r%nb), &
      & fgam(nF, r%nb), frc(nF, r%nb))
    call vlm_check(ctx, vlm_rotor_get_wake(ctx, r%dev, VLM_WAKE_C, c_loc(nodes), c_loc(le), &
      & c_loc(gam), c_loc(rc4), c_loc(ffc), c_loc(fgam), c_loc(frc)), 'get_wake')
    skew = 0._dp
    do ib = 1, r%nbConvect
      do icol = 1, r%ns
        do irow = r%rowNear, r%nNwake
          if (abs(gam(icol, irow, ib)) <= eps) cycle
          if (irow == r%rowNear) then
            p(:, 1) = le(:, icol, ib); p(:, 4) = le(:, icol + 1, ib)
          else
            p(:, 1) = nodes(:, icol, irow - 1, ib); p(:, 4) = nodes(:, icol + 1, irow - 1, ib)
          endif
          p(:, 2) = nodes(:, icol, irow, ib); p(:, 3) = nodes(:, icol + 1, irow, ib)
          x1 = p(:, 3) + p(:, 4) - p(:, 1) - p(:, 2)
          x2 = p(:, 4) + p(:, 1) - p(:, 2) - p(:, 3)
          skew(irow, icol, ib) = abs(dot_product(x1, x2)/ &
            & sqrt(dot_product(x1, x1)*dot_product(x2, x2)))
        enddo
      enddo
    enddo
    do ib = r%nbConvect + 1, r%nb
      skew(:, :, ib) = skew(:, :, 1)
    enddo
    do ib = 1, r%nb
      write (bid, '(I0.2)') ib
      open (newunit=u, file=trim(dir)//'/Results/r'//r%id//'b'//bid//'skew'//timestamp// &
        & '.csv', status='replace', action='write')
      write (u, '(2(A15))') 'max', 'avg'
      do irow = r%nNwake, r%rowNear, -1
        write (u, '(2(E15.7))') maxval(skew(irow, :, ib)), sum(skew(irow, :, ib))/r%ns
      enddo
      close (u)
    enddo
  end subroutine skew_file

  subroutine filaments_file(ctx, dir, timestamp, rotor, nr)
    !! filaments2file (libPostprocess.f90:363-473), the input of gridgen:
    !! every wing ring, every near-wake ring (all nNwake rows), the oldest
    !! near row's trailing filaments (vf(2), with -Γ) and the active far-wake
    !! filaments, as the reference's unformatted records of its vr_class and
    !! vf_class (a filament: fc(3,2), l0, lc, rVc0, rVc, age, ageAzimuthal; a
    !! ring: 4 filaments, gam, skew).  gridgen reads only fc, rVc and gam;
    !! the lengths, ages and skew are written as 0.
    type(c_ptr), intent(in) :: ctx
    character(len=*), intent(in) :: dir, timestamp
    type(rotor_class), intent(in) :: rotor(:)
    integer, intent(in) :: nr
    type :: dump_vf
      real(dp) :: fc(3, 2) = 0._dp
      real(dp) :: l0 = 0._dp, lc = 0._dp, rVc0 = 0._dp, rVc = 0._dp
      real(dp) :: age = 0._dp, ageAzimuthal = 0._dp
    end type dump_vf
    type :: dump_vr
      type(dump_vf) :: vf(4)
      real(dp) :: gam = 0._dp, skew = 0._dp
    end type dump_vr
    type(dump_vr), allocatable :: vrWing(:), vrNwake(:)
    type(dump_vf), allocatable :: vfNwakeTE(:), vfFwake(:)
    real(dp), allocatable :: gamNwakeTE(:), gamFwake(:)
    real(c_double), allocatable, target :: nodes(:, :, :, :), le(:, :, :), gam(:, :, :), &
      & rc4(:, :, :, :), ffc(:, :, :, :), fgam(:, :), frc(:, :)
    integer :: nW, nN, nT, nF, ir, ib, irow, icol, k, iw, in, it, jf, u
    real(dp) :: c(3, 4)

    do ir = 1, nr
      if (rotor(ir)%rowFar > rotor(ir)%nFwake) &
        & call host_fail('ERROR: Use filaments2file() only after development of far wake')
    enddo
    nW = 0; nN = 0; nT = 0; nF = 0
    do ir = 1, nr
      nW = nW + rotor(ir)%nb*rotor(ir)%nc*rotor(ir)%ns
      nN = nN + rotor(ir)%nb*rotor(ir)%nNwake*rotor(ir)%ns
      nT = nT + rotor(ir)%nb*rotor(ir)%ns
      nF = nF + (rotor(ir)%nFwake - rotor(ir)%rowFar + 1)*rotor(ir)%nb
    enddo
    allocate (vrWing(nW), vrNwake(nN), vfNwakeTE(nT), vfFwake(nF), gamNwakeTE(nT), gamFwake(nF))
    iw = 0; in = 0; it = 0; jf = 0
    do ir = 1, nr
      associate (r => rotor(ir))
        do ib = 1, r%nb
          do icol = 1, r%ns
            do irow = 1, r%nc
              iw = iw + 1
              do k = 1, 4
                vrWing(iw)%vf(k)%fc(:, 1) = r%blade(ib)%wiP(irow, icol)%vr%fc(:, k)
                vrWing(iw)%vf(k)%fc(:, 2) = r%blade(ib)%wiP(irow, icol)%vr%fc(:, mod(k, 4) + 1)
                vrWing(iw)%vf(k)%rVc = r%blade(ib)%wiP(irow, icol)%vr%rVc(k)
              enddo
              vrWing(iw)%gam = r%blade(ib)%wiP(irow, icol)%vr%gam
            enddo
          enddo
        enddo
        allocate (nodes(3, r%ns + 1, r%nNwake, r%nb), le(3, r%ns + 1, r%nb), &
          & gam(r%ns, r%nNwake, r%nb), rc4(4, r%ns, r%nNwake, r%nb), &
          & ffc(3, 2, r%nFwake, r%nb), fgam(r%nFwake, r%nb), frc(r%nFwake, r%nb))
        call vlm_check(ctx, vlm_rotor_get_wake(ctx, r%dev, VLM_WAKE_C, c_loc(nodes), &
          & c_loc(le), c_loc(gam), c_loc(rc4), c_loc(ffc), c_loc(fgam), c_loc(frc)), 'get_wake')
        do ib = 1, r%nb
          do icol = 1, r%ns
            do irow = 1, r%nNwake
              in = in + 1
              if (irow < r%rowNear) cycle
              ! ring corners: c1 and c4 on the upstream row (the shed line
              ! on the newest row), c2 and c3 on this row
              if (irow == r%rowNear) then
                c(:, 1) = le(:, icol, ib); c(:, 4) = le(:, icol + 1, ib)
              else
                c(:, 1) = nodes(:, icol, irow - 1, ib); c(:, 4) = nodes(:, icol + 1, irow - 1, ib)
              endif
              c(:, 2) = nodes(:, icol, irow, ib); c(:, 3) = nodes(:, icol + 1, irow, ib)
              do k = 1, 4
                vrNwake(in)%vf(k)%fc(:, 1) = c(:, k)
                vrNwake(in)%vf(k)%fc(:, 2) = c(:, mod(k, 4) + 1)
                vrNwake(in)%vf(k)%rVc = rc4(k, icol, irow, ib)
              enddo
              vrNwake(in)%gam = gam(icol, irow, ib)
            enddo
          enddo
        enddo
        do ib = 1, r%nb
          do icol = 1, r%ns
            it = it + 1
            vfNwakeTE(it)%fc(:, 1) = nodes(:, icol, r%nNwake, ib)
            vfNwakeTE(it)%fc(:, 2) = nodes(:, icol + 1, r%nNwake, ib)
            vfNwakeTE(it)%rVc = rc4(2, icol, r%nNwake, ib)
            gamNwakeTE(it) = gam(icol, r%nNwake, ib)*(-1._dp)
          enddo
        enddo
        do ib = 1, r%nb
          do irow = r%rowFar, r%nFwake
            jf = jf + 1
            vfFwake(jf)%fc = ffc(:, :, irow, ib)
            vfFwake(jf)%rVc = frc(irow, ib)
            gamFwake(jf) = fgam(irow, ib)
          enddo
        enddo
        deallocate (nodes, le, gam, rc4, ffc, fgam, frc)
      end associate
    enddo
    open (newunit=u, file=trim(dir)//'/Results/filaments'//timestamp//'.dat', &
      & status='replace', action='write', form='unformatted')
    write (u) nW
    write (u) nN
    write (u) nT
    write (u) nF
    write (u) vrWing, vrNwake
    write (u) vfNwakeTE, gamNwakeTE
    write (u) vfFwake, gamFwake
    close (u)
  end subroutine filaments_file

  subroutine inflow_launch(ctx, rotor, nr, ir, wakeSuppress)
    !! Enqueues inflow2file's velocities (libPostprocess.f90:900-946) at the
    !! section CPs of rotor ir: every rotor's wing less its bound vortices,
    !! plus its wake; they reach rotor(ir)%inflowV at vlm_wait
    type(c_ptr), intent(in) :: ctx
    type(rotor_class), intent(inout), target :: rotor(:)
    integer, intent(in) :: nr, ir, wakeSuppress
    integer(c_int) :: tir(3*nr), tset(3*nr), tfl(3*nr)
    real(c_double) :: p(3, rotor(ir)%ns, rotor(ir)%nb)
    integer :: jr, k, ib
    associate (r => rotor(ir))
      if (.not. allocated(r%inflowV)) allocate (r%inflowV(3, r%ns, r%nb))
      do ib = 1, r%nb
        p(:, :, ib) = r%blade(ib)%secCP
      enddo
      ! -(sum of bound vortices), then the wings and the wakes
      k = 0
      do jr = 1, nr
        k = k + 1; tir(k) = rotor(jr)%dev; tset(k) = VLM_SET_WING_BOUND; tfl(k) = 0
      enddo
      do jr = 1, nr
        k = k + 1; tir(k) = rotor(jr)%dev; tset(k) = VLM_SET_WING; tfl(k) = 0
        if (jr == 1) tfl(k) = VLM_TERM_NEGATE
      enddo
      do jr = 1, nr
        if (rotor(jr)%nNwake > 0 .and. wakeSuppress == 0) then
          k = k + 1; tir(k) = rotor(jr)%dev; tset(k) = VLM_SET_WAKE; tfl(k) = 0
        endif
      enddo
      call vlm_check(ctx, vlm_induced_velocity_terms_deferred(ctx, int(k, c_int), tir, tset, &
        & tfl, int(r%ns*r%nb, c_int64_t), p, c_loc(r%inflowV), 0_c_int), 'inflow')
    end associate
  end subroutine inflow_launch

  subroutine inflow_file(dir, timestamp, r)
    !! inflow2file's rXXbYYinflowDistNNNNN.curve: the inflow along -z at each
    !! section CP.  The span column is secCP . (hubCoords - yAxis), as the
    !! reference computes it.
    character(len=*), intent(in) :: dir, timestamp
    type(rotor_class), intent(in) :: r
    character(len=2) :: bid
    integer :: u, ib, is
    do ib = 1, r%nb
      write (bid, '(I0.2)') ib
      open (newunit=u, file=trim(dir)//'/Results/r'//r%id//'b'//bid//'inflowDist'// &
        & timestamp//'.curve', status='replace', action='write')
      write (u, '(2(A15))') 'secSpan', 'inflowVel'
      do is = 1, r%ns
        write (u, '(2(F15.7))') dot_product(r%blade(ib)%secCP(:, is), r%hubCoords - &
          & r%blade(ib)%yAxis), -r%inflowV(3, is, ib)
      enddo
      close (u)
    enddo
  end subroutine inflow_file

  subroutine gamma_file(dir, timestamp, r, first)
    !! gamma2file (libPostprocess.f90:949-978): the spanwise bound
    !! circulation -sum over the chord of Γ per blade.  The reference calls
    !! it with rotor 1 for every rotor whose switch fires (main.f90:777-783)
    !! and appends, so a step's later calls add copies to the same files.
    character(len=*), intent(in) :: dir, timestamp
    type(rotor_class), intent(in) :: r
    logical, intent(in) :: first
    character(len=2) :: bid
    integer :: u, ib, is
    do ib = 1, r%nb
      write (bid, '(I0.2)') ib
      if (first) then
        open (newunit=u, file=trim(dir)//'/Results/r'//r%id//'b'//bid//'gammaDist'// &
          & timestamp//'.curve', status='replace', action='write')
      else
        open (newunit=u, file=trim(dir)//'/Results/r'//r%id//'b'//bid//'gammaDist'// &
          & timestamp//'.curve', position='append', action='write')
      endif
      write (u, *) '# Row01'
      do is = 1, r%ns
        write (u, *) dot_product(r%blade(ib)%wiP(1, is)%CP - r%hubCoords, r%blade(ib)%yAxis), &
          & -1._dp*sum(r%blade(ib)%wiP(:, is)%vr%gam)
      enddo
      close (u)
    enddo
  end subroutine gamma_file

  subroutine surface_file(dir, r)
    !! geomSurface2file, libPostprocess.f90:475-615: the blade geometry after
    !! the initial pitch, one rXXbYYSurface.tec per blade (panel corners,
    !! vortex-ring corners, CPs with nCap, tauCapChord and tauCapSpan, and
    !! the sectional CPs with their normal and tangents)
    character(len=*), intent(in) :: dir
    type(rotor_class), intent(in) :: r
    real(dp) :: mesh(3, r%nc + 1, r%ns + 1)
    character(len=5) :: nxChar, nyChar
    character(len=2) :: bid
    integer :: u, ib, i, j, nx, ny, k

    do ib = 1, r%nb
      write (bid, '(I0.2)') ib
      open (newunit=u, file=trim(dir)//'/Results/r'//r%id//'b'//bid//'Surface.tec', &
        & status='replace', action='write')
      write (u, *) 'Title="r'//r%id//'b'//bid//'"'
      write (u, *) 'VARIABLES = "X" "Y" "Z" "nx" "ny" "nz"'
      nx = r%nc; ny = r%ns
      write (nxChar, '(I5)') nx + 1
      write (nyChar, '(I5)') ny + 1
      do k = 1, 2
        ! k = 1: panel corners PC; k = 2: corners of the vortex rings
        do j = 1, ny
          do i = 1, nx
            mesh(:, i, j) = corner(r%blade(ib)%wiP(i, j), 1, k)
          enddo
        enddo
        do i = 1, nx
          mesh(:, i, ny + 1) = corner(r%blade(ib)%wiP(i, ny), 4, k)
        enddo
        do j = 1, ny
          mesh(:, nx + 1, j) = corner(r%blade(ib)%wiP(nx, j), 2, k)
        enddo
        mesh(:, nx + 1, ny + 1) = corner(r%blade(ib)%wiP(nx, ny), 3, k)
        if (k == 1) then
          write (u, *) 'Zone I='//trim(nxChar)//' J='//trim(nyChar)//' K=1 T="PC"'
          write (u, *) 'DATAPACKING=BLOCK'
        else
          write (u, *) 'Zone I='//trim(nxChar)//' J='//trim(nyChar)//' K=1 T="VR"'
          write (u, *) 'DATAPACKING=BLOCK'
          write (u, *) 'VARSHARELIST=([4-6]=1)'
        endif
        write (u, *) ((mesh(1, i, j), i=1, nx + 1), j=1, ny + 1)
        write (u, *) ((mesh(2, i, j), i=1, nx + 1), j=1, ny + 1)
        write (u, *) ((mesh(3, i, j), i=1, nx + 1), j=1, ny + 1)
        if (k == 1) then
          write (u, *) ((0, i=1, nx + 1), j=1, ny + 1)
          write (u, *) ((0, i=1, nx + 1), j=1, ny + 1)
          write (u, *) ((0, i=1, nx + 1), j=1, ny + 1)
        endif
      enddo
      write (nxChar, '(I5)') nx
      write (nyChar, '(I5)') ny
      write (u, *) 'Zone I='//trim(nxChar)//' J='//trim(nyChar)//' K=1  T="CP, nCap"'
      write (u, *) 'DATAPACKING=BLOCK'
      do k = 1, 3
        write (u, *) ((r%blade(ib)%wiP(i, j)%CP(k), i=1, nx), j=1, ny)
      enddo
      do k = 1, 3
        write (u, *) ((r%blade(ib)%wiP(i, j)%nCap(k), i=1, nx), j=1, ny)
      enddo
      ! zones 4 and 5 carry the same title in the reference, and the titles
      ! continued over two source lines there carry two spaces before T=
      write (u, *) 'Zone I='//trim(nxChar)//' J='//trim(nyChar)//' K=1  T="CP, tauCapChord"'
      write (u, *) 'DATAPACKING=BLOCK'
      write (u, *) 'VARSHARELIST=([1-3]=3)'
      do k = 1, 3
        write (u, *) ((r%blade(ib)%wiP(i, j)%tauCapChord(k), i=1, nx), j=1, ny)
      enddo
      write (u, *) 'Zone I='//trim(nxChar)//' J='//trim(nyChar)//' K=1  T="CP, tauCapChord"'
      write (u, *) 'DATAPACKING=BLOCK'
      write (u, *) 'VARSHARELIST=([1-3]=3)'
      do k = 1, 3
        write (u, *) ((r%blade(ib)%wiP(i, j)%tauCapSpan(k), i=1, nx), j=1, ny)
      enddo
      write (nxChar, '(I5)') ny
      write (u, *) 'Zone I='//trim(nxChar)//' J=1 K=1 T="secCP, secNcap"'
      write (u, *) 'DATAPACKING=BLOCK'
      do k = 1, 3
        write (u, *) (r%blade(ib)%secCP(k, i), i=1, ny)
      enddo
      do k = 1, 3
        write (u, *) (r%blade(ib)%secNormalVec(k, i), i=1, ny)
      enddo
      write (u, *) 'Zone I='//trim(nxChar)//' J=1 K=1  T="secCP, secTauCapChord"'
      write (u, *) 'DATAPACKING=BLOCK'
      write (u, *) 'VARSHARELIST=([1-3]=6)'
      do k = 1, 3
        write (u, *) (r%blade(ib)%secTauCapChord(k, i), i=1, ny)
      enddo
      write (u, *) 'Zone I='//trim(nxChar)//' J=1 K=1  T="secCP, secTauCapSpan"'
      write (u, *) 'DATAPACKING=BLOCK'
      write (u, *) 'VARSHARELIST=([1-3]=6)'
      do k = 1, 3
        write (u, *) (r%blade(ib)%secTauCapSpan(k, i), i=1, ny)
      enddo
      close (u)
    enddo
  contains
    pure function corner(w, c, k) result(x)
      type(wingpanel_class), intent(in) :: w
      integer, intent(in) :: c, k
      real(dp) :: x(3)
      if (k == 1) then
        x = w%PC(:, c)
      else
        x = w%vr%fc(:, c)
      endif
    end function corner
  end subroutine surface_file

  subroutine age_azimuthal(r, dt)
    !! The azimuthal part of rotor_age_wake (classdef.f90:4343-4352): every
    !! active near-wake row and far-wake slot ages by dt*omegaSlow.
    type(rotor_class), intent(inout) :: r
    real(dp), intent(in) :: dt
    if (r%nNwake > 0) r%ageAzN(r%rowNear:r%nNwake) = r%ageAzN(r%rowNear:r%nNwake) + dt*r%omegaSlow
    if (r%nFwake > 0) r%ageAzF(r%rowFar:r%nFwake) = r%ageAzF(r%rowFar:r%nFwake) + dt*r%omegaSlow
  end subroutine age_azimuthal

  subroutine shift_azimuthal(r, rowFar)
    !! The azimuthal ages through rollup: a full far wake shifts by one slot
    !! (shiftFwake, classdef.f90:4575-4580 and 4500-4513), then the near wake
    !! by one row (shiftwake, 4604 and 4484-4498).  Neither shift resets the
    !! first slot's azimuthal age (only age), and the rolled-up filament keeps
    !! its slot's.
    type(rotor_class), intent(inout) :: r
    integer(c_int), intent(in) :: rowFar
    integer :: n
    n = r%nFwake
    if (n > 0 .and. rowFar == 1) r%ageAzF(2:n) = r%ageAzF(1:n - 1)
    n = r%nNwake
    r%ageAzN(2:n) = r%ageAzN(1:n - 1)
  end subroutine shift_azimuthal

  subroutine tip_file(ctx, dir, timestamp, r)
    !! tip2file, libPostprocess.f90:617-784: per blade the wing mesh with -Γ
    !! and azimuthal age (0 on the wing), the Γ-weighted centroid of corner 4
    !! over the rollup columns of every active near-wake row (corner 3 for
    !! the oldest row's far edge) with the row's extreme Γ, and the far-wake
    !! tip line, as rXXtipNNNNN.tec; the tip points also as rXXtipXYZNNNNN.csv.
    !! The wake comes from the device (vlm_rotor_get_wake).
    type(c_ptr), intent(in) :: ctx
    character(len=*), intent(in) :: dir, timestamp
    type(rotor_class), intent(in) :: r
    real(c_double), allocatable, target :: nodes(:, :, :, :), le(:, :, :), gam(:, :, :), &
      & rc4(:, :, :, :), ffc(:, :, :, :), fgam(:, :), frc(:, :)
    real(dp) :: wingMesh(3, r%nc + 1, r%ns + 1), nWakeTip(3, r%nNwake + 1)
    real(dp) :: gamRollup(r%nNwake), fWakeTip(3, max(r%nFwake, 1) + 1), gamSum, c(3)
    character(len=5) :: nxChar, nyChar
    integer :: u, u2, ib, i, j, nx, ny, nN, nF
    logical :: lowest

    nN = r%nNwake; nF = r%nFwake
    allocate (nodes(3, r%ns + 1, nN, r%nb), le(3, r%ns + 1, r%nb), gam(r%ns, nN, r%nb), &
      & rc4(4, r%ns, nN, r%nb), ffc(3, 2, max(nF, 1), r%nb), fgam(max(nF, 1), r%nb), &
      & frc(max(nF, 1), r%nb))
    call vlm_check(ctx, vlm_rotor_get_wake(ctx, r%dev, VLM_WAKE_C, c_loc(nodes), c_loc(le), &
      & c_loc(gam), c_loc(rc4), c_loc(ffc), c_loc(fgam), c_loc(frc)), 'get_wake')
    ! Γ is negative for positive Omega*pitch, so the extreme is the lowest
    lowest = sign(1._dp, r%Omega*r%controlPitch(1)) > eps
    open (newunit=u, file=trim(dir)//'/Results/r'//r%id//'tip'//timestamp//'.tec', &
      & status='replace', action='write')
    open (newunit=u2, file=trim(dir)//'/Results/r'//r%id//'tipXYZ'//timestamp//'.csv', &
      & status='replace', action='write')
    write (u, *) 'Title = "Wing and Tip"'
    write (u, *) 'VARIABLES = "X" "Y" "Z" "GAM" "ageAzimuthal"'
    do ib = 1, r%nb
      nx = r%nc; ny = r%ns
      write (nxChar, '(I5)') nx + 1
      write (nyChar, '(I5)') ny + 1
      do j = 1, ny
        do i = 1, nx
          wingMesh(:, i, j) = r%blade(ib)%wiP(i, j)%PC(:, 1)
        enddo
      enddo
      do i = 1, nx
        wingMesh(:, i, ny + 1) = r%blade(ib)%wiP(i, ny)%PC(:, 4)
      enddo
      do j = 1, ny
        wingMesh(:, nx + 1, j) = r%blade(ib)%wiP(nx, j)%PC(:, 2)
      enddo
      wingMesh(:, nx + 1, ny + 1) = r%blade(ib)%wiP(nx, ny)%PC(:, 3)
      write (u, *) 'Zone I='//trim(nxChar)//' J='//trim(nyChar)//' K=1  T="Blade"'
      write (u, *) 'DATAPACKING=BLOCK'
      write (u, *) 'VARLOCATION=([4,5]=CELLCENTERED)'
      write (u, *) ((wingMesh(1, i, j), i=1, nx + 1), j=1, ny + 1)
      write (u, *) ((wingMesh(2, i, j), i=1, nx + 1), j=1, ny + 1)
      write (u, *) ((wingMesh(3, i, j), i=1, nx + 1), j=1, ny + 1)
      write (u, *) ((-1._dp*r%blade(ib)%wiP(i, j)%vr%gam, i=1, nx), j=1, ny)
      write (u, *) ((0._dp*radToDeg, i=1, nx), j=1, ny)
      ! near wake: corner 4 of ring (i, j) is LE node j+1 of row i, the shed
      ! line on the newest row and row i-1's trailing nodes behind it
      nx = nN
      write (nxChar, '(I5)') nx - (r%rowNear - 1) + 1
      do i = r%rowNear, nx
        gamRollup(i) = gam(ny, i, ib)
        gamSum = 0._dp
        nWakeTip(:, i) = 0._dp
        do j = r%rollupStart, r%rollupEnd
          nWakeTip(:, i) = nWakeTip(:, i) + corner4(i, j)*gam(j, i, ib)
          gamSum = gamSum + gam(j, i, ib)
          if (lowest) then
            if (gam(j, i, ib) < gamRollup(i)) gamRollup(i) = gam(j, i, ib)
          else
            if (gam(j, i, ib) > gamRollup(i)) gamRollup(i) = gam(j, i, ib)
          endif
        enddo
        if (abs(gamSum) > eps) then
          nWakeTip(:, i) = nWakeTip(:, i)/gamSum
        else
          nWakeTip(:, i) = corner4(i, r%rollupEnd)
        endif
      enddo
      ! the oldest row's corner 3 (trailing node j+1)
      gamSum = 0._dp
      nWakeTip(:, nx + 1) = 0._dp
      do j = r%rollupStart, r%rollupEnd
        nWakeTip(:, nx + 1) = nWakeTip(:, nx + 1) + nodes(:, j + 1, nx, ib)*gam(j, nx, ib)
        gamSum = gamSum + gam(j, nx, ib)
      enddo
      if (abs(gamSum) > eps) then
        nWakeTip(:, nx + 1) = nWakeTip(:, nx + 1)/gamSum
      else
        nWakeTip(:, nx + 1) = nodes(:, r%rollupEnd + 1, nx, ib)
      endif
      write (u, *) 'Zone I='//trim(nxChar)//' J=1    K=1  T="NearWake"'
      write (u, *) 'DATAPACKING=BLOCK'
      write (u, *) 'VARLOCATION=([4,5]=CELLCENTERED)'
      write (u, *) (nWakeTip(1, i), i=r%rowNear, nx + 1)
      write (u, *) (nWakeTip(2, i), i=r%rowNear, nx + 1)
      write (u, *) (nWakeTip(3, i), i=r%rowNear, nx + 1)
      write (u, *) (-1._dp*gamRollup(i), i=r%rowNear, nx)
      write (u, *) (r%ageAzN(i)*radToDeg, i=r%rowNear, nx)
      write (u2, *) 'X        Y         Z'
      do i = r%rowNear, nx + 1
        write (u2, *) nWakeTip(1, i), nWakeTip(2, i), nWakeTip(3, i)
      enddo
      ! far wake: the LE end of every active filament and the TE end of the
      ! oldest, or the oldest near ring's corner 3 twice when there is none
      if (r%rowFar <= nF) then
        nx = nF
        write (nxChar, '(I5)') nx - (r%rowFar - 1) + 1
        do i = r%rowFar, nx
          fWakeTip(:, i) = ffc(:, 2, i, ib)
        enddo
        fWakeTip(:, nx + 1) = ffc(:, 1, nF, ib)
        write (u, *) 'Zone I='//trim(nxChar)//' J=1   K=1   T="FarWake"'
        write (u, *) 'DATAPACKING=BLOCK'
        write (u, *) 'VARLOCATION=([4,5]=CELLCENTERED)'
        write (u, *) (fWakeTip(1, i), i=r%rowFar, nx + 1)
        write (u, *) (fWakeTip(2, i), i=r%rowFar, nx + 1)
        write (u, *) (fWakeTip(3, i), i=r%rowFar, nx + 1)
        write (u, *) (-1._dp*fgam(i, ib), i=r%rowFar, nx)
        write (u, *) (r%ageAzF(i)*radToDeg, i=r%rowFar, nx)
        do i = r%rowFar, nx + 1
          write (u2, *) fWakeTip(1, i), fWakeTip(2, i), fWakeTip(3, i)
        enddo
      else
        c = nodes(:, ny + 1, nN, ib)
        write (nxChar, '(I5)') 2
        write (u, *) 'Zone I='//trim(nxChar)//' J=1   K=1   T="FarWake"'
        write (u, *) 'DATAPACKING=BLOCK'
        write (u, *) 'VARLOCATION=([4,5]=CELLCENTERED)'
        write (u, *) c(1), c(1)
        write (u, *) c(2), c(2)
        write (u, *) c(3), c(3)
        write (u, *) 0._dp
        write (u, *) 0._dp
      endif
    enddo
    close (u)
    close (u2)

  contains

    function corner4(i, j) result(p)
      integer, intent(in) :: i, j
      real(dp) :: p(3)
      if (i == r%rowNear) then
        p = le(:, j + 1, ib)
      else
        p = nodes(:, j + 1, i - 1, ib)
      endif
    end function corner4

  end subroutine tip_file

  subroutine geom_file(ctx, dir, timestamp, r, wakeSuppress)
    !! geom2file, libPostprocess.f90:143-331, for lifting surfaces: per
    !! blade the wing panel corners and -Γ, the active near-wake lattice and
    !! -Γ, the far-wake tip line and -Γ and, once built, the prescribed far
    !! wake with -Γ and the core radius, as Tecplot blocks in the
    !! reference's list-directed format.  The wake
    !! comes from the device (vlm_rotor_get_wake).  The skew column is 0:
    !! the reference fills the wake rings' skew only for skew plots.
    type(c_ptr), intent(in) :: ctx
    character(len=*), intent(in) :: dir, timestamp
    type(rotor_class), intent(in) :: r
    integer, intent(in) :: wakeSuppress
    real(c_double), allocatable, target :: nodes(:, :, :, :), le(:, :, :), gam(:, :, :), &
      & rc4(:, :, :, :), ffc(:, :, :, :), fgam(:, :), frc(:, :)
    real(dp) :: wingMesh(3, r%nc + 1, r%ns + 1), wakeMesh(3, r%nNwake + 1, r%ns + 1)
    real(dp) :: wakeTip(3, max(r%nFwake, 1) + 1), tip1(3)
    character(len=5) :: nxChar, nyChar
    character(len=2) :: bid
    integer, parameter :: PF_N = 240
    real(c_double), allocatable, target :: pfc(:, :, :, :), pgam(:, :), prc(:, :)
    integer(c_int) :: pbuilt
    integer :: u, ib, i, j, nx, ny, nN, nF
    logical :: wake

    nN = r%nNwake; nF = r%nFwake
    wake = wakeSuppress == 0 .and. nN > 0
    pbuilt = 0
    if (wake) then
      allocate (nodes(3, r%ns + 1, nN, r%nb), le(3, r%ns + 1, r%nb), gam(r%ns, nN, r%nb), &
        & rc4(4, r%ns, nN, r%nb), ffc(3, 2, max(nF, 1), r%nb), fgam(max(nF, 1), r%nb), &
        & frc(max(nF, 1), r%nb))
      call vlm_check(ctx, vlm_rotor_get_wake(ctx, r%dev, VLM_WAKE_C, c_loc(nodes), c_loc(le), &
        & c_loc(gam), c_loc(rc4), c_loc(ffc), c_loc(fgam), c_loc(frc)), 'get_wake')
      allocate (pfc(3, 2, PF_N, r%nb), pgam(PF_N, r%nb), prc(PF_N, r%nb))
      call vlm_check(ctx, vlm_rotor_get_prescribed_wake(ctx, r%dev, VLM_WAKE_C, c_loc(pfc), &
        & c_loc(pgam), c_loc(prc), pbuilt), 'get_prescribed_wake')
    endif
    ! the reference appends (position='append'); within one run each file is
    ! written once, so replacing gives its output also when a case is rerun
    open (newunit=u, file=trim(dir)//'/Results/r'//r%id//'wingNwake'//timestamp//'.tec', &
      & status='replace', action='write')
    write (u, *) 'Title = "Wing and Wake"'
    write (u, *) 'VARIABLES = "X" "Y" "Z" "GAM" "skew"'
    do ib = 1, r%nb
      write (bid, '(I0.2)') ib
      nx = r%nc; ny = r%ns
      write (nxChar, '(I5)') nx + 1
      write (nyChar, '(I5)') ny + 1
      do j = 1, ny
        do i = 1, nx
          wingMesh(:, i, j) = r%blade(ib)%wiP(i, j)%PC(:, 1)
        enddo
      enddo
      do i = 1, nx
        wingMesh(:, i, ny + 1) = r%blade(ib)%wiP(i, ny)%PC(:, 4)
      enddo
      do j = 1, ny
        wingMesh(:, nx + 1, j) = r%blade(ib)%wiP(nx, j)%PC(:, 2)
      enddo
      wingMesh(:, nx + 1, ny + 1) = r%blade(ib)%wiP(nx, ny)%PC(:, 3)
      write (u, *) 'Zone I='//trim(nxChar)//' J='//trim(nyChar)//' K=1  T="Blade '//bid//'"'
      write (u, *) 'DATAPACKING=BLOCK'
      write (u, *) 'VARLOCATION=([4]=CELLCENTERED,[5]=CELLCENTERED)'
      write (u, *) ((wingMesh(1, i, j), i=1, nx + 1), j=1, ny + 1)
      write (u, *) ((wingMesh(2, i, j), i=1, nx + 1), j=1, ny + 1)
      write (u, *) ((wingMesh(3, i, j), i=1, nx + 1), j=1, ny + 1)
      write (u, *) ((-1._dp*r%blade(ib)%wiP(i, j)%vr%gam, i=1, nx), j=1, ny)
      write (u, *) ((r%blade(ib)%wiP(i, j)%vr%gam*0._dp, i=1, nx), j=1, ny)
      if (.not. wake) cycle
      ! near wake: corner 1 of the active rings (the shed line on the newest
      ! row), the last column's corner 4, the oldest row's corners 2 and 3
      nx = nN
      write (nxChar, '(I5)') nx - (r%rowNear - 1) + 1
      write (nyChar, '(I5)') ny + 1
      do j = 1, ny + 1
        wakeMesh(:, r%rowNear, j) = le(:, j, ib)
        do i = r%rowNear + 1, nx + 1
          wakeMesh(:, i, j) = nodes(:, j, i - 1, ib)
        enddo
      enddo
      write (u, *) 'Zone I='//trim(nxChar)//' J='//trim(nyChar)//' K=1  T="NearWake"'
      write (u, *) 'DATAPACKING=BLOCK'
      write (u, *) 'VARLOCATION=([4]=CELLCENTERED, [5]=CELLCENTERED)'
      write (u, *) ((wakeMesh(1, i, j), i=r%rowNear, nx + 1), j=1, ny + 1)
      write (u, *) ((wakeMesh(2, i, j), i=r%rowNear, nx + 1), j=1, ny + 1)
      write (u, *) ((wakeMesh(3, i, j), i=r%rowNear, nx + 1), j=1, ny + 1)
      write (u, *) ((-1._dp*gam(j, i, ib), i=r%rowNear, nx), j=1, ny)
      write (u, *) ((0._dp, i=r%rowNear, nx), j=1, ny)
      ! far wake: the LE end of every active filament and the TE end of the
      ! oldest, or the oldest near ring's corner 3 twice when there is none
      tip1 = nodes(:, ny + 1, nN, ib)
      if (r%rowFar <= nF) then
        nx = nF
        write (nxChar, '(I5)') nx - (r%rowFar - 1) + 1
        do i = r%rowFar, nx
          wakeTip(:, i) = ffc(:, 2, i, ib)
        enddo
        wakeTip(:, nx + 1) = ffc(:, 1, nF, ib)
        write (u, *) 'Zone I='//trim(nxChar)//' J=1   K=1   T="FarWake"'
        write (u, *) 'DATAPACKING=BLOCK'
        write (u, *) 'VARLOCATION=([4]=CELLCENTERED ,[5]=CELLCENTERED)'
        write (u, *) (wakeTip(1, i), i=r%rowFar, nx + 1)
        write (u, *) (wakeTip(2, i), i=r%rowFar, nx + 1)
        write (u, *) (wakeTip(3, i), i=r%rowFar, nx + 1)
        write (u, *) (-1._dp*fgam(i, ib), i=r%rowFar, nx)
        write (u, *) (frc(i, ib)*0._dp, i=r%rowFar, nx)
      else
        write (nxChar, '(I5)') 2
        write (u, *) 'Zone I='//trim(nxChar)//' J=1   K=1   T="FarWake"'
        write (u, *) 'DATAPACKING=BLOCK'
        write (u, *) 'VARLOCATION=([4]=CELLCENTERED, [5]=CELLCENTERED)'
        write (u, *) tip1(1), tip1(1)
        write (u, *) tip1(2), tip1(2)
        write (u, *) tip1(3), tip1(3)
        write (u, *) 0._dp
        write (u, *) 0._dp
      endif
      if (pbuilt /= 0) then
        ! prescribed far wake: the LE end of every filament and the TE end
        ! of the last, -Γ and the core radius
        write (nxChar, '(I5)') PF_N + 1
        write (u, *) 'Zone I='//trim(nxChar)//' J=1   K=1   T="PrescFarWake"'
        write (u, *) 'DATAPACKING=BLOCK'
        write (u, *) 'VARLOCATION=([4]=CELLCENTERED ,[5]=CELLCENTERED)'
        write (u, *) (pfc(1, 2, i, ib), i=1, PF_N), pfc(1, 1, PF_N, ib)
        write (u, *) (pfc(2, 2, i, ib), i=1, PF_N), pfc(2, 1, PF_N, ib)
        write (u, *) (pfc(3, 2, i, ib), i=1, PF_N), pfc(3, 1, PF_N, ib)
        write (u, *) (-1._dp*pgam(i, ib), i=1, PF_N)
        write (u, *) (prc(i, ib), i=1, PF_N)
      else
        ! no prescribed far wake: the same redundant point
        write (nxChar, '(I5)') 2
        write (u, *) 'Zone I='//trim(nxChar)//' J=1   K=1   T="PrescFarWake"'
        write (u, *) 'DATAPACKING=BLOCK'
        write (u, *) 'VARLOCATION=([4]=CELLCENTERED, [5]=CELLCENTERED)'
        write (u, *) tip1(1), tip1(1)
        write (u, *) tip1(2), tip1(2)
        write (u, *) tip1(3), tip1(3)
        write (u, *) 0._dp
        write (u, *) 0._dp
      endif
    enddo
    close (u)
  end subroutine geom_file

  subroutine sub_iterate(ctx, rotor, nr, nsub, with_wake, running_max)
    !! Sub-iterations of the RHS and the solve (main.f90:120-211 initial,
    !! 522-615 every step): after a deferred compute_rhs_and_solve, wait,
    !! map Γ, and solve again until max over rotors of
    !! norm2(gamVec - gamVecPrev) <= eps, at most nsub more times.  The
    !! per-step test of the reference exits inside the rotor loop as soon as
    !! the running maximum is below eps (running_max); the initial test
    !! takes the maximum over all rotors.
    type(c_ptr), intent(in) :: ctx
    type(rotor_class), intent(inout), target :: rotor(:)
    integer, intent(in) :: nr, nsub
    logical, intent(in) :: with_wake, running_max
    integer :: i, ir
    real(dp) :: res
    do i = 0, nsub
      call vlm_check(ctx, vlm_wait(ctx), 'wait')
      call finish_solve(rotor, nr)
      res = 0._dp
      do ir = 1, nr
        res = max(res, norm2(rotor(ir)%gamVec - rotor(ir)%gamVecPrev))
        if (running_max .and. res <= eps) return
      enddo
      if (res <= eps) return
      if (i < nsub) call compute_rhs_and_solve(ctx, rotor, nr, with_wake)
    enddo
  end subroutine sub_iterate

  subroutine forces_launch(ctx, rotor, nr, ir, density, dt)
    !! Enqueues the loads of rotor ir on the device (vlm_rotor_loads): the
    !! load velocities at the collocation points, the wing's own rings less
    !! every rotor's bound vortices (classdef.f90:1376-1418, 1734), then
    !! the sectional and rotor loads of forceCalcSwitch = 0; the results
    !! reach rotor(ir)%lout at vlm_wait and compute_forces unpacks them.
    type(c_ptr), intent(in) :: ctx
    type(rotor_class), intent(inout), target :: rotor(:)
    integer, intent(in) :: nr, ir
    real(dp), intent(in) :: density, dt
    integer :: jr
    integer(c_int) :: tir(nr + 1), tset(nr + 1), tfl(nr + 1)

    associate (r => rotor(ir))
      if (.not. allocated(r%lout)) allocate (r%lout(19*r%nbConvect*r%ns + 9))
      do jr = 1, nr
        tir(jr) = rotor(jr)%dev; tset(jr) = VLM_SET_WING_BOUND; tfl(jr) = 0
      enddo
      tir(nr + 1) = r%dev; tset(nr + 1) = VLM_SET_WING; tfl(nr + 1) = VLM_TERM_NEGATE
      call vlm_check(ctx, vlm_rotor_loads(ctx, r%dev, int(nr + 1, c_int), tir, tset, tfl, &
        & density, dt, r%Omega, int(r%spanwiseLiftSwitch, c_int), c_loc(r%lout)), 'loads')
    end associate
  end subroutine forces_launch

  subroutine compute_forces(rotor, ir)
    !! After vlm_wait: the device loads of forces_launch into the blade
    !! sections and the rotor totals that the force files and CT read
    !! (records of vlm_rotor_loads)
    type(rotor_class), intent(inout) :: rotor(:)
    integer, intent(in) :: ir
    integer :: ib, is, k

    associate (r => rotor(ir))
      k = 0
      do ib = 1, r%nbConvect
        associate (b => r%blade(ib))
          do is = 1, r%ns
            b%secChordwiseResVel(:, is) = r%lout(k + 1:k + 3)
            b%secAlpha(is) = r%lout(k + 4)
            b%secPhi(is) = r%lout(k + 5)
            b%secViz(is) = r%lout(k + 6)
            b%secVix(is) = r%lout(k + 7)
            b%secTheta(is) = r%lout(k + 8)
            b%secLift(:, is) = r%lout(k + 9:k + 11)
            b%secLiftInPlane(:, is) = r%lout(k + 12:k + 14)
            b%secLiftOutPlane(:, is) = r%lout(k + 15:k + 17)
            b%secCL(is) = r%lout(k + 18)
            b%secCLu(is) = r%lout(k + 19)
            k = k + 19
          enddo
        end associate
      enddo
      r%forceInertial = r%lout(k + 1:k + 3)
      r%lift = r%lout(k + 4:k + 6)
      r%liftUnsteady = r%lout(k + 7:k + 9)
      r%drag = 0._dp
    end associate
  end subroutine compute_forces

  function rotor_CT(r) result(ct)
    !! Non-dimensional lift (CL or CT), libPostprocess.f90:824-828
    type(rotor_class), intent(in) :: r
    real(dp) :: ct
    ct = sign(1._dp, dot_product(r%lift, r%zAxisBody))*norm2(r%lift)/r%nonDimforceDenominator
  end function rotor_CT

  subroutine force_line(u, r, it)
    integer, intent(in) :: u, it
    type(rotor_class), intent(in) :: r
    character(len=5) :: ts
    real(dp) :: d
    write (ts, '(I0.5)') it
    d = r%nonDimforceDenominator
    write (u, '(A,9(E15.7))') ts, rotor_CT(r), norm2(r%drag)/d, norm2(r%liftUnsteady)/d, &
      & 0._dp, 0._dp, 0._dp, r%forceInertial(1)/d, r%forceInertial(2)/d, r%forceInertial(3)/d
  end subroutine force_line

  subroutine force_files(uf, ud, dir, r, it)
    !! force2file, libPostprocess.f90:814-882, at every force step: the
    !! ForceNonDim and ForceDim rows and the ForceDist table of every
    !! convected blade
    integer, intent(in) :: uf, ud, it
    character(len=*), intent(in) :: dir
    type(rotor_class), intent(in) :: r
    character(len=5) :: ts
    character(len=2) :: bc2
    integer :: ib
    call force_line(uf, r, it)
    write (ts, '(I0.5)') it
    write (ud, '(A,11(E15.7))') ts, norm2(r%lift), norm2(r%drag), r%lift, r%drag, &
      & r%forceInertial
    do ib = 1, r%nbConvect
      write (bc2, '(I0.2)') ib
      call dist_file(trim(dir)//'/Results/r'//r%id//'b'//bc2//'ForceDist'//ts//'.csv', &
        & r%blade(ib), r%hubCoords, r%ns)
    enddo
  end subroutine force_files

  subroutine dist_file(path, b, hub, ns)
    !! Sectional loads of one blade in the format of force2file's ForceDist
    !! table (libPostprocess.f90:848-881).  Drag is zero in this load model
    !! (forceCalcSwitch = 0: no induced or profile drag, classdef.f90:1856-1863),
    !! so secCD and secDrag are written as 0.
    character(len=*), intent(in) :: path
    type(blade_class), intent(in) :: b
    real(dp), intent(in) :: hub(3)
    integer, intent(in) :: ns
    integer :: u, is
    open (newunit=u, file=path, status='replace', action='write')
    write (u, '(16(A15))') 'secSpan', 'secCL', 'secCD', 'secCLu', 'secLift', 'secDrag', &
      & 'secLiftInPl', 'secLiftOutPl', 'secArea', 'secVel', 'secChord', 'secTheta', &
      & 'secAlpha', 'secPhi', 'secViz', 'secVix'
    do is = 1, ns
      write (u, '(16(E15.7))') dot_product(b%secCP(:, is) - hub, b%yAxis), b%secCL(is), &
        & 0._dp, b%secCLu(is), norm2(b%secLift(:, is)), 0._dp, &
        & norm2(b%secLiftInPlane(:, is)), norm2(b%secLiftOutPlane(:, is)), b%secArea(is), &
        & norm2(b%secChordwiseResVel(:, is)), b%secChord(is), b%secTheta(is)*radToDeg, &
        & b%secAlpha(is)*radToDeg, b%secPhi(is)*radToDeg, b%secViz(is), b%secVix(is)
    enddo
    close (u)
  end subroutine dist_file

  subroutine params_file(path, r, nt, dt, nr, density, velSound, sw)
    !! rXXParams.json with the keys and value conventions of params2file
    !! (libPostprocess.f90:52-141): angles in degrees except phi/theta/psi,
    !! counts after the init-time conversions.  Written as a JSON object the
    !! reference's tests/parseResults.py getParams reads.
    character(len=*), intent(in) :: path
    type(rotor_class), intent(in) :: r
    integer, intent(in) :: nt, nr
    real(dp), intent(in) :: dt, density, velSound
    type(switches_class), intent(in) :: sw
    integer :: u
    open (newunit=u, file=path, status='replace', action='write')
    write (u, '(A)') '{'
    call ji('nt', nt); call jr('dt', dt); call ji('nr', nr)
    call ji('restartWriteNt', sw%restartWriteNt); call ji('restartFromNt', sw%restartFromNt)
    call ji('ntSub', sw%ntSub); call ji('ntSubInit', sw%ntSubInit)
    call jr('density', density); call jr('velSound', velSound)
    call ji('wakePlot', sw%wakePlot); call ji('wakeTipPlot', sw%wakeTipPlot)
    call ji('rotorForcePlot', sw%rotorForcePlot); call ji('gridPlot', sw%gridPlot)
    call ji('wakeDissipation', sw%wakeDissipation); call ji('wakeStrain', sw%wakeStrain)
    call ji('wakeBurst', sw%wakeBurst); call ji('slowStart', sw%slowStart)
    call ji('slowStartNt', sw%slowStartNt); call ji('fdScheme', sw%fdScheme)
    call ji('initWakeVelNt', sw%initWakeVelNt); call ji('probe', sw%probe)
    call ji('nb', r%nb); call ji('propConvention', r%propConvention)
    write (u, '(A)') ' "geometryFile": "'//r%geometryFile//'",'
    call ji('spanSpacing', r%spanSpacing); call ji('chordSpacing', r%chordSpacing)
    call ji('nCamberFiles', 0); call ji('surfaceType', r%surfaceType)
    call ji('imagePlane', r%imagePlane); call ji('imageRotorNum', r%imageRotorNum)
    call ji('nc', r%nc); call ji('ns', r%ns); call ji('nNwake', r%nNwake)
    call jr('radius', r%radius); call jr('root_cut', r%root_cut); call jr('chord', r%chord)
    call jr('preconeAngle', r%preconeAngle*radToDeg); call jr('Omega', r%Omega)
    call jr('phi', r%pts(1)); call jr('theta', r%pts(2)); call jr('psi', r%pts(3))
    call jr('theta0', r%controlPitch(1)*radToDeg); call jr('thetaC', r%controlPitch(2)*radToDeg)
    call jr('thetaS', r%controlPitch(3)*radToDeg); call jr('thetaTwist', r%thetaTwist*radToDeg)
    call jr('u', r%velBody(1)); call jr('v', r%velBody(2)); call jr('w', r%velBody(3))
    call jr('nonDimForceDenom', r%nonDimforceDenominator); call jr('alpha0', 0._dp)
    call ji('wakeTruncateNt', r%wakeTruncateNt); call ji('prescWakeNt', r%prescWakeNt)
    call jr('initWakeVel', r%initWakeVel); call jr('psiStart', r%psiStart*radToDeg)
    call jr('forceCalcSwitch', r%forceCalcSwitch*radToDeg)
    call jr('apparentViscCoeff', r%apparentViscCoeff); call jr('decayCoeff', r%decayCoeff)
    call ji('axisymmetrySwitch', r%axisymmetrySwitch); call ji('bladeDynamicsSwitch', 0)
    call jr('flapInitial', r%flapInitial); call jr('dFlapInitial', r%dflapInitial)
    call jr('Iflap', r%Iflap); call jr('cflap', r%cflap); call jr('kflap', r%kflap)
    call jr('MflapConstant', r%MflapConstant); call ji('pitchDynamicsSwitch', 0)
    write (u, '(A,ES24.16E3)') ' "dpitch": ', r%dpitch
    write (u, '(A)') '}'
    close (u)
  contains
    subroutine ji(k, v)
      character(len=*), intent(in) :: k
      integer, intent(in) :: v
      write (u, '(A,I0,A)') ' "'//k//'": ', v, ','
    end subroutine ji
    subroutine jr(k, v)
      character(len=*), intent(in) :: k
      real(dp), intent(in) :: v
      write (u, '(A,ES24.16E3,A)') ' "'//k//'": ', v, ','
    end subroutine jr
  end subroutine params_file

  ! ------------------------------------------------------------------
  ! The time loop, main.f90:1-1459
  ! ------------------------------------------------------------------
  subroutine run_case(dir, precision, nt_override, verbose, ct_hist, ct_cap, hist_len, stats, &
      & tail_steps, predictor, pred_tol)
    character(len=*), intent(in) :: dir
    integer, intent(in) :: precision, nt_override, verbose, ct_cap, hist_len, tail_steps
    integer, intent(in) :: predictor  !! bit 0: learned pass 2; bit 1: predicted-wake
                                      !! loop over every row (reference quirk off)
    real(dp), intent(in) :: pred_tol
    real(dp), intent(inout) :: ct_hist(ct_cap, hist_len)
    type(volcanor_stats), intent(out) :: stats
    type(switches_class) :: sw
    type(rotor_class), allocatable, target :: rotor(:)
    type(c_ptr) :: ctx
    integer :: nt, nr, ir, ib, iter, rn, rf, uforce(16), udim(16), nforce
    real(dp) :: dt, density, velSound, kinematicVisc, t
    character(len=2) :: rc2, bc2
    character(len=5) :: ts5
    type(vlm_rotor_desc) :: d
    integer(c_int) :: rowNear, rowFar, dev
    integer(c_int64_t) :: c0, c1, c2, crate
    logical :: write_out, loads, probing
    integer :: ntStart, nprobe, uprobe, k
    real(dp), allocatable :: probe(:, :), probeVel(:, :)
    real(c_double), allocatable, target :: probeLoc(:, :), probeOut(:, :)
    integer(c_int), allocatable :: pir(:), pset(:), pfl(:)

    call system_clock(c0, crate)
    call read_config(dir, nt, dt, nr, density, velSound, kinematicVisc, sw)
    if (sw%slowStart < 0 .or. sw%slowStart > 3) &
      & call host_fail('ERROR: slowStart must be 0..3 (body dynamics are not supported)')
    if (sw%slowStart > 0 .and. sw%slowStartNt <= 0) call host_fail('ERROR: slowStartNt must be positive')
    if (sw%fdScheme < 0 .or. sw%fdScheme > 5) call host_fail('ERROR: fdScheme must be 0..5')
    if (sw%ntSub < 0 .or. sw%ntSubInit < 0) call host_fail('ERROR: ntSub and ntSubInit must be >= 0')
    if (sw%restartFromNt < 0 .or. sw%restartWriteNt < 0) &
      & call host_fail('ERROR: restartFromNt and restartWriteNt must be >= 0')
    if (sw%restartFromNt > 0 .and. iand(predictor, 1) /= 0) &
      & call host_fail('ERROR: a restart does not carry the learned predictor; run without it')
    if (nr > 16) call host_fail('ERROR: at most 16 rotors')
    ! velocity probes (main.f90:83-93): read only for probe = 1, used at every
    ! step that is a multiple of probe; any other value would leave the
    ! reference's probe arrays unallocated
    nprobe = 0
    if (sw%probe /= 0) then
      if (sw%probe /= 1) call host_fail('ERROR: probe must be 0 or 1')
      call open_old(uprobe, trim(dir)//'/probes.in')
      read (uprobe, *) nprobe
      allocate (probe(3, nprobe), probeVel(3, nprobe), probeLoc(3, nprobe), probeOut(3, nprobe))
      do k = 1, nprobe
        read (uprobe, *) probe(:, k), probeVel(:, k)
      enddo
      close (uprobe)
    endif
    allocate (rotor(nr))
    do ir = 1, nr
      write (rc2, '(I0.2)') ir
      call rotor_readGeom(rotor(ir), dir, 'geom'//rc2//'.nml')
    enddo
    do ir = 1, nr
      if (rotor(ir)%surfaceType < 0) then
        ! the reference initialises an image from its source's initialised
        ! state (main.f90:32-38), so the source comes first and is lifting
        if (rotor(ir)%imageRotorNum < 1 .or. rotor(ir)%imageRotorNum >= ir) &
          & call host_fail('ERROR: imageRotorNum must name a rotor before the image')
        if (rotor(rotor(ir)%imageRotorNum)%surfaceType < 0) &
          & call host_fail('ERROR: the source of an image rotor must not be an image')
        call rotor_init(rotor(ir), ir, density, dt, nt, sw, dir, rotor(rotor(ir)%imageRotorNum))
      else
        call rotor_init(rotor(ir), ir, density, dt, nt, sw, dir)
      endif
    enddo
    if (nt_override > 0) nt = nt_override
    write_out = sw%rotorForcePlot /= 0 .and. verbose >= 0

    call vlm_check(c_null_ptr, vlm_create(ctx, 0_c_int), 'vlm_create')
    active_ctx = ctx
    call vlm_check(ctx, vlm_set_precision(ctx, int(precision, c_int)), 'set_precision')
    if (iand(predictor, 2) /= 0) call vlm_check(ctx, vlm_set_bug_compat(ctx, 0_c_int), &
      & 'set_bug_compat')
    if (iand(predictor, 1) /= 0) call vlm_check(ctx, vlm_set_predictor(ctx, 1_c_int, pred_tol, 32_c_int, &
      & 8_c_int, 4_c_int, rotor(1)%chord), 'set_predictor')

    do ir = 1, nr
      associate (r => rotor(ir))
        d%nb = r%nb; d%nc = r%nc; d%ns = r%ns; d%nNwake = r%nNwake; d%nFwake = r%nFwake
        d%nbConvect = r%nbConvect; d%rollupStart = r%rollupStart; d%rollupEnd = r%rollupEnd
        d%rollupMin = 0
        if (sign(1._dp, r%Omega*r%controlPitch(1)) > eps) d%rollupMin = 1
        d%suppressFwake = r%suppressFwakeSwitch
        call vlm_check(ctx, vlm_rotor_create(ctx, d, r%spanwiseCore, r%streamwiseCoreVec, dev), &
          & 'rotor_create')
        r%dev = dev
        call vlm_check(ctx, vlm_rotor_set_init_wake_vel(ctx, r%dev, r%initWakeVel, &
          & int(sw%initWakeVelNt, c_int)), 'set_init_wake_vel')
        ! initial pitch, main.f90:43-58; an image takes its source's
        ! (negated for the plane z = 0)
        block
          integer :: ib
          do ib = 1, r%nb
            if (r%surfaceType < 0) then
              r%blade(ib)%theta = rotor(r%imageRotorNum)%blade(ib)%theta
              if (r%imagePlane == 3) r%blade(ib)%theta = -1._dp*r%blade(ib)%theta
            else
              r%blade(ib)%theta = rotor_gettheta(r, r%psiStart, ib)
            endif
            call blade_rot_pitch(r%blade(ib), sign(1._dp, r%Omega)*r%blade(ib)%theta)
          enddo
        end block
        call rotor_upload_wing(ctx, r)
        if (r%surfaceType < 0) then
          ! no influence matrix: the image's circulation is its source's
          ! negated (main.f90:66-81, 189-193)
          call vlm_check(ctx, vlm_rotor_set_image(ctx, r%dev, rotor(r%imageRotorNum)%dev, &
            & int(r%imagePlane, c_int)), 'set_image')
        else
          block
            real(c_double), allocatable :: cp(:, :, :, :), ncap(:, :, :, :)
            allocate (cp(3, r%nc, r%ns, r%nb), ncap(3, r%nc, r%ns, r%nb))
            call rotor_cp_arrays(r, cp, ncap)
            call vlm_check(ctx, vlm_rotor_calc_aic(ctx, r%dev, cp, ncap, c_null_ptr), 'calcAIC')
          end block
        endif
      end associate
    enddo

    ! initial solution without wake; a slow start begins at rest (main.f90:97-105);
    ! a custom trajectory starts from its first line (112-117)
    do ir = 1, nr
      if (rotor(ir)%customTrajectorySwitch == 1) then
        rotor(ir)%velBody = rotor(ir)%velBodyHistory(:, 1)
        rotor(ir)%omegaBody = rotor(ir)%omegaBodyHistory(:, 1)
      endif
      rotor(ir)%omegaSlow = rotor(ir)%Omega
      if (sw%slowStart > 0) rotor(ir)%omegaSlow = 0._dp
    enddo
    t = 0._dp
    call compute_rhs_and_solve(ctx, rotor, nr, .false.)
    if (sw%ntSubInit /= 0) call sub_iterate(ctx, rotor, nr, sw%ntSubInit, .false., .false.)
    call vlm_check(ctx, vlm_wait(ctx), 'wait')
    call finish_solve(rotor, nr)
    do ir = 1, nr
      call vlm_check(ctx, vlm_rotor_init_cursors(ctx, rotor(ir)%dev), 'init_cursors')
      if (rotor(ir)%nNwake > 0) &
        & call vlm_check(ctx, vlm_rotor_assignshed(ctx, rotor(ir)%dev, VLM_EDGE_TE), 'assignshed TE')
    enddo
    nforce = 0
    if (write_out) then
      block
        integer(c_int) :: mk
        ! an existing directory is fine; the opens below report any failure
        mk = c_mkdir(trim(dir)//'/Results'//c_null_char, int(o'755', c_int))
      end block
      do ir = 1, nr
        write (rc2, '(I0.2)') ir
        call surface_file(dir, rotor(ir))
        call params_file(trim(dir)//'/Results/r'//rc2//'Params.json', rotor(ir), nt, dt, nr, &
          & density, velSound, sw)
        open (newunit=uforce(ir), file=trim(dir)//'/Results/r'//rc2//'ForceNonDim.csv', &
          & status='replace', action='write')
        write (uforce(ir), '(A5,9(A15))') 'iter', 'CL/CT', 'CD/CQ', 'CLu', 'CDi', 'CD0', 'CDu', &
          & 'CFx', 'CFy', 'CFz'
        ! init_plots, libPostprocess.f90:10-49
        open (newunit=udim(ir), file=trim(dir)//'/Results/r'//rc2//'ForceDim.csv', &
          & status='replace', action='write')
        write (udim(ir), '(A5,11(A15))') 'iter', 'LiftMag', 'DragMag', 'Lx', 'Ly', 'Lz', &
          & 'Dx', 'Dy', 'Dz', 'FInertx', 'FInerty', 'FInertz'
        block
          integer :: ud
          open (newunit=ud, file=trim(dir)//'/Results/r'//rc2//'bladedynamics.csv', &
            & status='replace', action='write')
          write (ud, '(A5,5(A15))') 'iter', 'pitch', 'flap', 'dflap', 'w', 'omega'
          close (ud)
        end block
      enddo
    endif
    if (sw%rotorForcePlot /= 0) then
      do ir = 1, nr
        call forces_launch(ctx, rotor, nr, ir, density, dt)
      enddo
      call vlm_check(ctx, vlm_wait(ctx), 'wait')
      do ir = 1, nr
        call compute_forces(rotor, ir)
        if (ir <= ct_cap .and. hist_len >= 1) ct_hist(ir, 1) = rotor_CT(rotor(ir))
        if (write_out) call force_files(uforce(ir), udim(ir), dir, rotor(ir), 0)
      enddo
    endif

    call system_clock(c1)
    c2 = c1
    ntStart = 1
    ! resume from a restart file (main.f90:390-398): the state at the end of
    ! step restartFromNt, read after the initial solution as there
    if (sw%restartFromNt > 0) then
      block
        integer :: ur
        write (ts5, '(I0.5)') sw%restartFromNt
        call open_old(ur, trim(dir)//'/Restart/restart'//ts5//'.dat', unformatted=.true.)
        call restart_io(ctx, ur, rotor, nr, t, .false.)
        close (ur)
      end block
      ntStart = sw%restartFromNt + 1
    endif
    do iter = ntStart, nt
      if (iter == nt - tail_steps + 1) then
        call vlm_check(ctx, vlm_synchronize(ctx), 'synchronize')
        if (tail_steps > 0) then
          call vlm_check(ctx, vlm_profile_enable(ctx, 1_c_int), 'profile_enable')
          call vlm_check(ctx, vlm_profile_reset(ctx), 'profile_reset')
        endif
        call system_clock(c2)
      endif
      t = t + dt
      if (verbose > 0) then
        if (mod(iter, verbose) == 0) print '(A,I0,A,I0)', ' step ', iter, ' of ', nt
      endif
      do ir = 1, nr
        call vlm_check(ctx, vlm_rotor_begin_step(ctx, rotor(ir)%dev, int(iter, c_int)), 'begin_step')
        call vlm_check(ctx, vlm_rotor_cursors(ctx, rotor(ir)%dev, rowNear, rowFar), 'cursors')
        rotor(ir)%rowNear = rowNear; rotor(ir)%rowFar = rowFar
        if (rotor(ir)%customTrajectorySwitch == 1) then   ! main.f90:419-426
          rotor(ir)%velBody = rotor(ir)%velBodyHistory(:, iter)
          rotor(ir)%omegaBody = rotor(ir)%omegaBodyHistory(:, iter)
        endif
        rotor(ir)%omegaSlow = slow_omega(sw, iter, rotor(ir)%Omega)
      enddo
      do ir = 1, nr
        call rotor_move(rotor(ir), rotor(ir)%velBody*dt)
        call rotor_rot_pts(rotor(ir), rotor(ir)%omegaBody*dt, rotor(ir)%cgCoords)
        call rotor_rot_advance(rotor(ir), rotor(ir)%omegaSlow*dt, .false.)
        call rotor_upload_wing(ctx, rotor(ir))
      enddo
      if (sw%wakeSuppress == 0) then
        do ir = 1, nr
          if (rotor(ir)%nNwake <= 0) cycle
          call vlm_check(ctx, vlm_rotor_assignshed(ctx, rotor(ir)%dev, VLM_EDGE_LE), 'assignshed LE')
          call vlm_check(ctx, vlm_rotor_age_wake(ctx, rotor(ir)%dev, dt), 'age_wake')
          call age_azimuthal(rotor(ir), dt)
          if (sw%wakeDissipation == 1) call vlm_check(ctx, vlm_rotor_dissipate_wake(ctx, &
            & rotor(ir)%dev, dt, kinematicVisc, rotor(ir)%apparentViscCoeff, rotor(ir)%decayCoeff), &
            & 'dissipate_wake')
          ! far-wake burst every wakeBurst steps (main.f90:490-496)
          if (sw%wakeBurst /= 0) then
            if (mod(iter, sw%wakeBurst) == 0) call vlm_check(ctx, vlm_rotor_burst_wake(ctx, &
              & rotor(ir)%dev, rotor(ir)%skewLimit, rotor(ir)%chord), 'burst_wake')
          endif
          ! near-wake skew plot (main.f90:497-502)
          if (write_out .and. rotor(ir)%skewPlotSwitch /= 0) then
            if (mod(iter, rotor(ir)%skewPlotSwitch) == 0) then
              write (ts5, '(I0.5)') iter
              call skew_file(ctx, dir, ts5, rotor(ir))
            endif
          endif
        enddo
      endif
      ! wing and wake plot, main.f90:509-513
      if (write_out .and. sw%wakePlot /= 0) then
        if (mod(iter, sw%wakePlot) == 0) then
          write (ts5, '(I0.5)') iter
          do ir = 1, nr
            call geom_file(ctx, dir, ts5, rotor(ir), sw%wakeSuppress)
          enddo
        endif
      endif
      ! wake tip plot, main.f90:515-518
      if (write_out .and. sw%wakeTipPlot /= 0 .and. sw%wakeSuppress == 0) then
        if (mod(iter, sw%wakeTipPlot) == 0) then
          write (ts5, '(I0.5)') iter
          do ir = 1, nr
            if (rotor(ir)%nNwake > 0) call tip_file(ctx, dir, ts5, rotor(ir))
          enddo
        endif
      endif
      call compute_rhs_and_solve(ctx, rotor, nr, sw%wakeSuppress == 0)
      if (sw%ntSub /= 0) &
        & call sub_iterate(ctx, rotor, nr, sw%ntSub, sw%wakeSuppress == 0, .true.)
      loads = .false.
      if (sw%rotorForcePlot /= 0) loads = mod(iter, sw%rotorForcePlot) == 0
      if (loads) then
        do ir = 1, nr
          call forces_launch(ctx, rotor, nr, ir, density, dt)
        enddo
      endif
      ! velocity probes (probes2file, libPostprocess.f90:333-361): the probe's
      ! own velocity plus every rotor's wing and wake velocity at its moving
      ! position, on the wake before convection (main.f90:792-797)
      probing = .false.
      if (sw%probe /= 0 .and. nprobe > 0) probing = mod(iter, sw%probe) == 0
      if (probing) then
        do k = 1, nprobe
          probeLoc(:, k) = probe(:, k) + probeVel(:, k)*t
        enddo
        probeOut = probeVel
        if (.not. allocated(pir)) allocate (pir(2*nr), pset(2*nr), pfl(2*nr))
        k = 0
        do ir = 1, nr
          k = k + 1; pir(k) = rotor(ir)%dev; pset(k) = VLM_SET_WING; pfl(k) = 0
          if (rotor(ir)%nNwake > 0 .and. sw%wakeSuppress == 0) then
            k = k + 1; pir(k) = rotor(ir)%dev; pset(k) = VLM_SET_WAKE; pfl(k) = 0
          endif
        enddo
        call vlm_check(ctx, vlm_induced_velocity_terms_deferred(ctx, int(k, c_int), pir, &
          & pset, pfl, int(nprobe, c_int64_t), probeLoc, c_loc(probeOut), 1_c_int), 'probes')
      endif
      ! filament dumps for gridgen (main.f90:785-790): the wake before
      ! convection with this step's circulation
      if (write_out .and. sw%gridPlot /= 0) then
        if (mod(iter, sw%gridPlot) == 0) then
          call vlm_check(ctx, vlm_wait(ctx), 'wait')
          call finish_solve(rotor, nr)
          write (ts5, '(I0.5)') iter
          call filaments_file(ctx, dir, ts5, rotor, nr)
        endif
      endif
      ! inflow plots (main.f90:767-774), on the wake before convection
      do ir = 1, nr
        if (.not. write_out .or. rotor(ir)%inflowPlotSwitch == 0) cycle
        if (mod(iter, rotor(ir)%inflowPlotSwitch) /= 0) cycle
        call inflow_launch(ctx, rotor, nr, ir, sw%wakeSuppress)
      enddo
      ! The wake passes are enqueued before the host waits for the
      ! circulation and the load velocities; the loads (main.f90:632-719)
      ! depend only on the wing and its circulation, which the convection
      ! does not change, so they are computed while the passes run.
      ! the prescribed far wake's azimuth step is omegaSlow*dt of this step
      do ir = 1, nr
        if (rotor(ir)%prescWakeNt > 0) call vlm_check(ctx, vlm_rotor_set_prescribed_wake(ctx, &
          & rotor(ir)%dev, int(rotor(ir)%prescWakeNt, c_int), &
          & int(rotor(ir)%prescWakeGenNt, c_int), rotor(ir)%omegaSlow*dt), 'set_prescribed_wake')
      enddo
      if (sw%wakeSuppress == 0) &
        & call vlm_check(ctx, vlm_convect(ctx, int(sw%fdScheme, c_int), int(iter, c_int), dt), &
        & 'convect')
      call vlm_check(ctx, vlm_wait(ctx), 'wait')
      call finish_solve(rotor, nr)
      if (write_out) then
        write (ts5, '(I0.5)') iter
        do ir = 1, nr
          if (rotor(ir)%inflowPlotSwitch == 0) cycle
          if (mod(iter, rotor(ir)%inflowPlotSwitch) == 0) call inflow_file(dir, ts5, rotor(ir))
        enddo
        k = 0
        do ir = 1, nr
          if (rotor(ir)%gammaPlotSwitch == 0) cycle
          if (mod(iter, rotor(ir)%gammaPlotSwitch) /= 0) cycle
          call gamma_file(dir, ts5, rotor(1), k == 0)
          k = k + 1
        enddo
      endif
      if (probing .and. verbose >= 0) then
        write (ts5, '(I0.5)') iter
        open (newunit=uprobe, file=trim(dir)//'/Results/probes'//ts5//'.csv', &
          & status='replace', action='write')
        write (uprobe, '(6(A15))') 'u', 'v', 'w', 'x', 'y', 'z'
        do k = 1, nprobe
          write (uprobe, '(6(E15.7))') probeOut(:, k), probeLoc(:, k)
        enddo
        close (uprobe)
      endif
      if (loads) then
        do ir = 1, nr
          call compute_forces(rotor, ir)
          if (ir <= ct_cap .and. iter + 1 <= hist_len) ct_hist(ir, iter + 1) = rotor_CT(rotor(ir))
          if (write_out) call force_files(uforce(ir), udim(ir), dir, rotor(ir), iter)
        enddo
      endif
      if (sw%wakeSuppress == 0) then
        if (sw%wakeStrain == 1) then
          do ir = 1, nr
            if (rotor(ir)%nNwake > 0) &
              & call vlm_check(ctx, vlm_rotor_strain_wake(ctx, rotor(ir)%dev), 'strain_wake')
          enddo
        endif
        do ir = 1, nr
          if (rotor(ir)%nNwake <= 0) cycle
          if (rotor(ir)%surfaceType < 0) then
            ! an image copies its source's whole wake, mirrored, in place of
            ! the rollup and the shed (main.f90:1433-1436)
            call vlm_check(ctx, vlm_rotor_mirror_wake(ctx, rotor(ir)%dev, VLM_WAKE_A), 'mirror_wake')
            rotor(ir)%ageAzN = rotor(rotor(ir)%imageRotorNum)%ageAzN
            rotor(ir)%ageAzF = rotor(rotor(ir)%imageRotorNum)%ageAzF
            cycle
          endif
          call vlm_check(ctx, vlm_rotor_cursors(ctx, rotor(ir)%dev, rowNear, rowFar), 'cursors')
          if (rowNear == 1) then
            call vlm_check(ctx, vlm_rotor_rollup(ctx, rotor(ir)%dev), 'rollup')
            call shift_azimuthal(rotor(ir), rowFar)
          endif
          call vlm_check(ctx, vlm_rotor_assignshed(ctx, rotor(ir)%dev, VLM_EDGE_TE), 'assignshed TE')
        enddo
      endif
      ! restart file at the end of the step (main.f90:1442-1450)
      if (sw%restartWriteNt /= 0 .and. verbose >= 0) then
        if (mod(iter, sw%restartWriteNt) == 0) then
          block
            integer :: uw
            integer(c_int) :: mk
            ! an existing directory is fine; the open below reports any failure
            mk = c_mkdir(trim(dir)//'/Restart'//c_null_char, int(o'755', c_int))
            write (ts5, '(I0.5)') iter
            open (newunit=uw, file=trim(dir)//'/Restart/restart'//ts5//'.dat', &
              & status='replace', action='write', form='unformatted')
            call restart_io(ctx, uw, rotor, nr, t, .true.)
            close (uw)
          end block
        endif
      endif
    enddo
    call vlm_check(ctx, vlm_synchronize(ctx), 'synchronize')
    block
      integer(c_int64_t) :: cend
      call system_clock(cend)
      stats%wall_loop = real(cend - c1, dp)/real(crate, dp)
      stats%wall_steps_tail = real(cend - c2, dp)/real(crate, dp)
      stats%wall_total = real(cend - c0, dp)/real(crate, dp)
    end block
    stats%nt = nt
    stats%tail_steps = tail_steps
    stats%bs_kernel_ms = 0._dp; stats%bs_pairs = 0._dp; stats%bs_launches = 0
    stats%bs_pairs_all = 0._dp; stats%bs_launches_all = 0
    if (tail_steps > 0) then
      call vlm_check(ctx, vlm_profile_get(ctx, stats%bs_kernel_ms, stats%bs_launches, &
        & stats%bs_pairs), 'profile_get')
      call vlm_check(ctx, vlm_profile_get_all(ctx, stats%bs_pairs_all, stats%bs_launches_all), &
        & 'profile_get_all')
      call vlm_check(ctx, vlm_profile_enable(ctx, 0_c_int), 'profile_enable')
    endif
    block
      real(c_double) :: le, ll
      call vlm_check(ctx, vlm_predictor_stats(ctx, stats%pred_full, stats%pred_learned, &
        & stats%pred_fallbacks, le, ll), 'predictor_stats')
      stats%pred_last_err = le
      call vlm_check(ctx, vlm_predictor_errors(ctx, stats%pred_mlp_rms, stats%pred_base_rms, &
        & stats%pred_judged), 'predictor_errors')
    end block
    if (write_out) then
      do ir = 1, nr
        close (uforce(ir))
        close (udim(ir))
      enddo
    endif
    rn = 0; rf = 0
    active_ctx = c_null_ptr
    call vlm_check(ctx, vlm_destroy(ctx), 'vlm_destroy')
  end subroutine run_case

  subroutine volcanor_host_cleanup() bind(C, name='volcanor_host_cleanup')
    !! After a run stopped by an error (host_guard.c): destroys its libvlm
    !! context, which frees every device buffer of the run.
    integer(c_int) :: rc
    integer :: k, ios
    logical :: op
    if (c_associated(active_ctx)) rc = vlm_destroy(active_ctx)
    active_ctx = c_null_ptr
    ! files the stopped run left open (newunit numbers are negative)
    do k = -1, -4096, -1
      op = .false.
      inquire (unit=k, opened=op, iostat=ios)
      if (ios == 0 .and. op) close (k, iostat=ios)
    enddo
  end subroutine volcanor_host_cleanup

  subroutine volcanor_run_impl(cdir, precision, nt_override, verbose, ct_hist, ct_cap, &
      & hist_len, stats, tail_steps, predictor, pred_tol) bind(C, name='volcanor_run_impl')
    !! The whole solver for the C entry volcanor_run (host_guard.c): runs
    !! the case in directory cdir (null-terminated) and returns the CL/CT
    !! history ct_hist(ct_cap, nt+1).
    character(kind=c_char), intent(in) :: cdir(*)
    integer(c_int), value :: precision, nt_override, verbose, ct_cap, hist_len, tail_steps
    integer(c_int), value :: predictor
    real(c_double), value :: pred_tol
    real(c_double), intent(inout) :: ct_hist(ct_cap, hist_len)
    type(volcanor_stats), intent(out) :: stats
    character(len=1024) :: dir
    integer :: i
    dir = ''
    do i = 1, 1024
      if (cdir(i) == c_null_char) exit
      dir(i:i) = cdir(i)
    enddo
    call run_case(trim(dir), int(precision), int(nt_override), int(verbose), ct_hist, &
      & int(ct_cap), int(hist_len), stats, int(tail_steps), int(predictor), pred_tol)
  end subroutine volcanor_run_impl

end module volcanor_host
