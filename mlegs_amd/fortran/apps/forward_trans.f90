!> Forward transform tutorial (src/apps/forward_trans.f90): loads
!> <logdir>scalar_PPP.dat (written by backward_trans), transforms it to FFF
!> and writes <logdir>scalar_FFF.dat.
!> Usage: forward_trans [input_2d.params]
program forward_trans
  use mlegs_amd
  implicit none
  integer(i4) :: glb_sz(3)
  character(len=256) :: input_params_file = './input_2d.params'
  logical :: exists
  type(scalar) :: s
  if (command_argument_count() .ge. 1) call get_command_argument(1, input_params_file)
  write(*, *) 'Program Started'
  inquire(file=input_params_file, exist=exists)
  if (exists) call read_input(input_params_file)
  call tfm%init()
  write(*, *) 'Initialization'
  glb_sz = [tfm%nrdim, tfm%npdim, tfm%nzdim]
  call s%init(glb_sz, [1, 0, 2])
  call s%set_space('PPP')
  call mload(trim(adjustl(logdir)) // 'scalar_PPP.dat', s, is_binary=.false., is_global=.true.)
  call trans(s, 'FFF', tfm)
  call msave(s, trim(adjustl(logdir)) // 'scalar_FFF.dat', is_binary=.false., is_global=.true.)
  call s%dealloc()
  call tfm%dealloc()
  write(*, *) 'Program Finished'
end program
