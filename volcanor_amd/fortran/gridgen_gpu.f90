!! gridgen_gpu: the velocity field on a Cartesian grid from the filament dumps
!! of a run (src/gridgen.f90 of the reference), with the Biot-Savart sums on
!! the GPU (vlm_filament_vind, one launch per dump over every cell centre).
!! Run from a case directory: reads gridconfig.nml (template 0.2: nx, ny, nz,
!! xyzMin, xyzMax, vel, fileRangeStart/Step/End) and
!! Results/filamentsNNNNN.dat, writes Results/gridNNNNN.tec.
!!   gridgen_gpu [casedir]
program gridgen_gpu
  use iso_c_binding
  use vlm_c
  implicit none
  integer, parameter :: dp = kind(1.d0)
  type :: dump_vf
    real(dp) :: fc(3, 2), l0, lc, rVc0, rVc, age, ageAzimuthal
  end type dump_vf
  type :: dump_vr
    type(dump_vf) :: vf(4)
    real(dp) :: gam, skew
  end type dump_vr
  integer :: nx, ny, nz, fileRangeStart, fileRangeStep, fileRangeEnd
  real(dp) :: xyzMin(3), xyzMax(3), vel(3)
  character(len=10) :: fileFormatVersion
  character(len=1024) :: dir
  namelist /VERSION/ fileFormatVersion
  namelist /INPUTS/ nx, ny, nz, xyzMin, xyzMax, vel, fileRangeStart, fileRangeStep, fileRangeEnd
  real(dp), allocatable :: grid(:, :, :, :), centre(:, :, :, :), vc(:, :, :, :)
  real(dp), allocatable :: xv(:), yv(:), zv(:), a(:, :), b(:, :), g(:), rc(:)
  type(dump_vr), allocatable :: vrWing(:), vrNwake(:)
  type(dump_vf), allocatable :: vfNwakeTE(:), vfFwake(:)
  real(dp), allocatable :: gamNwakeTE(:), gamFwake(:)
  integer :: u, ix, iy, iz, f, k, n, nW, nN, nT, nF, nfil
  integer(c_int64_t) :: m
  character(len=5) :: nxc, nyc, nzc, ts
  type(c_ptr) :: ctx

  dir = '.'
  if (command_argument_count() >= 1) call get_command_argument(1, dir)
  open (newunit=u, file=trim(dir)//'/gridconfig.nml', status='old', action='read')
  read (u, nml=VERSION)
  if (adjustl(fileFormatVersion) /= '0.2') &
    & error stop 'ERROR: gridconfig.nml template version does not match'
  read (u, nml=INPUTS)
  close (u)
  if (xyzMin(1) > xyzMax(1) .or. xyzMin(2) > xyzMax(2) .or. xyzMin(3) > xyzMax(3)) &
    & error stop 'ERROR: All XYZmin values should be greater than XYZmax values'

  ! grid nodes (libMath.f90:138-157 linspace) and cell centres, as the
  ! reference forms them (gridgen.f90:60-85)
  allocate (grid(3, nx, ny, nz), centre(3, nx - 1, ny - 1, nz - 1), vc(3, nx - 1, ny - 1, nz - 1))
  xv = span(xyzMin(1), xyzMax(1), nx)
  yv = span(xyzMin(2), xyzMax(2), ny)
  zv = span(xyzMin(3), xyzMax(3), nz)
  do iz = 1, nz
    do iy = 1, ny
      do ix = 1, nx
        grid(:, ix, iy, iz) = [xv(ix), yv(iy), zv(iz)]
      enddo
    enddo
  enddo
  do iz = 1, nz - 1
    do iy = 1, ny - 1
      do ix = 1, nx - 1
        centre(:, ix, iy, iz) = grid(:, ix, iy, iz) + grid(:, ix + 1, iy, iz) &
          & + grid(:, ix + 1, iy + 1, iz) + grid(:, ix + 1, iy + 1, iz + 1) &
          & + grid(:, ix, iy + 1, iz) + grid(:, ix, iy + 1, iz + 1) &
          & + grid(:, ix, iy, iz + 1) + grid(:, ix + 1, iy, iz + 1)
      enddo
    enddo
  enddo
  centre = centre*0.125_dp
  m = int(nx - 1, c_int64_t)*(ny - 1)*(nz - 1)
  write (nxc, '(I5)') nx
  write (nyc, '(I5)') ny
  write (nzc, '(I5)') nz

  call vlm_check(ctx, vlm_create(ctx, 0_c_int), 'vlm_create')
  do f = fileRangeStart, fileRangeEnd, fileRangeStep
    write (ts, '(I0.5)') f
    open (newunit=u, file=trim(dir)//'/Results/filaments'//ts//'.dat', status='old', &
      & action='read', form='unformatted')
    read (u) nW
    read (u) nN
    read (u) nT
    read (u) nF
    allocate (vrWing(nW), vrNwake(nN), vfNwakeTE(nT), vfFwake(nF), gamNwakeTE(nT), gamFwake(nF))
    read (u) vrWing, vrNwake
    read (u) vfNwakeTE, gamNwakeTE
    read (u) vfFwake, gamFwake
    close (u)
    ! One explicit filament list.  The reference's wing loop assigns
    ! instead of accumulating (gridgen.f90:122), so of the wing only its
    ! last ring contributes; then every near-wake ring, the trailing
    ! filaments and the far wake.
    nfil = 4*(min(nW, 1) + nN) + nT + nF
    allocate (a(3, max(nfil, 1)), b(3, max(nfil, 1)), g(max(nfil, 1)), rc(max(nfil, 1)))
    n = 0
    if (nW > 0) call ring(vrWing(nW))
    do k = 1, nN
      call ring(vrNwake(k))
    enddo
    do k = 1, nT
      call fil(vfNwakeTE(k), gamNwakeTE(k))
    enddo
    do k = 1, nF
      call fil(vfFwake(k), gamFwake(k))
    enddo
    vc = 0._dp
    call vlm_check(ctx, vlm_filament_vind(ctx, int(n, c_int64_t), a, b, g, rc, m, centre, vc, &
      & 0_c_int), 'filament_vind')
    do k = 1, 3
      vc(k, :, :, :) = vc(k, :, :, :) + vel(k)
    enddo
    open (newunit=u, file=trim(dir)//'/Results/grid'//ts//'.tec', status='replace', &
      & action='write')
    write (u, *) 'TITLE = "Grid"'
    write (u, *) 'VARIABLES = "X" "Y" "Z" "U" "V" "W"'
    write (u, *) 'ZONE I='//trim(nxc)//' J='//trim(nyc)//' K='//trim(nzc)//' T="Data"'
    write (u, *) 'DATAPACKING=BLOCK'
    write (u, *) 'VARLOCATION=([4]=CELLCENTERED,[5]=CELLCENTERED,[6]=CELLCENTERED)'
    do k = 1, 3
      write (u, *) (((grid(k, ix, iy, iz), ix=1, nx), iy=1, ny), iz=1, nz)
    enddo
    do k = 1, 3
      write (u, *) (((vc(k, ix, iy, iz), ix=1, nx - 1), iy=1, ny - 1), iz=1, nz - 1)
    enddo
    close (u)
    deallocate (vrWing, vrNwake, vfNwakeTE, vfFwake, gamNwakeTE, gamFwake, a, b, g, rc)
  enddo
  call vlm_check(ctx, vlm_destroy(ctx), 'vlm_destroy')

contains

  function span(x0, x1, np) result(x)
    real(dp), intent(in) :: x0, x1
    integer, intent(in) :: np
    real(dp) :: x(np), dx
    integer :: i
    dx = (x1 - x0)/(np - 1)
    x = [((i*dx), i=0, np - 1)]
    x = x + x0
  end function span

  subroutine fil(v, gam)
    type(dump_vf), intent(in) :: v
    real(dp), intent(in) :: gam
    n = n + 1
    a(:, n) = v%fc(:, 1)
    b(:, n) = v%fc(:, 2)
    g(n) = gam
    rc(n) = v%rVc
  end subroutine fil

  subroutine ring(r)
    type(dump_vr), intent(in) :: r
    integer :: i
    do i = 1, 4
      call fil(r%vf(i), r%gam)
    enddo
  end subroutine ring

end program gridgen_gpu
