!> Inverse Laplacian tutorial (src/apps/inverse_laplacian.f90): loads
!> <logdir>scalar_FFF_Laplacian.dat (written by laplacian), applies idel2
!> and writes scalar_FFF_from_inv.dat.
!> Usage: inverse_laplacian [input_2d.params]
program inverse_laplacian
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
  call s%init(glb_sz, [0, 1, 2])
  call s%set_space('FFF')
  call mload(trim(adjustl(logdir)) // 'scalar_FFF_Laplacian.dat', s, is_binary=.false., is_global=.true.)
  call idel2(s, tfm)
  call msave(s, trim(adjustl(logdir)) // 'scalar_FFF_from_inv.dat', is_binary=.false., is_global=.true.)
  call s%dealloc()
  call tfm%dealloc()
  write(*, *) 'Program Finished'
end program
