!! volcanor_gpu: runs the case in the current directory (config.nml,
!! geomXX.nml) like the reference executable (src/main.f90), with every
!! Biot-Savart sum, the wake and the circulation solve on the GPU.
!!   volcanor_gpu [casedir] [fp64|mixed]
program volcanor_gpu
  use volcanor_host
  implicit none
  character(len=1024) :: dir, prec
  real(dp) :: hist(1, 1)
  type(volcanor_stats) :: st
  integer :: p

  dir = '.'
  prec = 'fp64'
  if (command_argument_count() >= 1) call get_command_argument(1, dir)
  if (command_argument_count() >= 2) call get_command_argument(2, prec)
  p = 0
  if (trim(prec) == 'mixed') p = 1
  call run_case(trim(dir), p, 0, 1, hist, 1, 1, st, 0, 0, 1.e-4_dp)
  print '(A,F10.3,A,I0,A)', ' wall ', st%wall_total, ' s for ', st%nt, ' steps'
end program volcanor_gpu
