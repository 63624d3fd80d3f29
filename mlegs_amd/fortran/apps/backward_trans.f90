!> Backward transform tutorial (src/apps/backward_trans.f90): two spectral
!> coefficients set to 1 (global entries (2,2,1) and (2,3,1) of the FFF
!> array), saved, transformed to PPP, saved.
!> Writes <logdir>scalar_FFF_original.dat, scalar_PPP.dat and coords_*.dat.
!> Usage: backward_trans [input_2d.params]
program backward_trans
  use mlegs_amd
  implicit none
  integer(i4) :: glb_sz(3)
  character(len=256) :: input_params_file = './input_2d.params'
  logical :: exists
  complex(p8), dimension(:,:,:), allocatable :: array_glb
  type(scalar) :: s
  if (command_argument_count() .ge. 1) call get_command_argument(1, input_params_file)
  write(*, *) 'Program Started'
  inquire(file=input_params_file, exist=exists)
  if (exists) call read_input(input_params_file)
  call tfm%init()
  call colloc_pts(tfm)
  write(*, *) 'Initialization'
  glb_sz = [tfm%nrdim, tfm%npdim, tfm%nzdim]
  allocate(array_glb(glb_sz(1), glb_sz(2), glb_sz(3)))
  array_glb = 0.0_p8
  array_glb(2, 2, 1) = 1.0_p8
  array_glb(2, 3, 1) = 1.0_p8
  call s%init(glb_sz, [0, 1, 2])
  call s%set_space('FFF')
  call s%disassemble(array_glb)
  call msave(s, trim(adjustl(logdir)) // 'scalar_FFF_original.dat', is_binary=.false., is_global=.true.)
  call trans(s, 'PPP', tfm)
  call msave(s, trim(adjustl(logdir)) // 'scalar_PPP.dat', is_binary=.false., is_global=.true.)
  deallocate(array_glb)
  call s%dealloc()
  call tfm%dealloc()
  write(*, *) 'Program Finished'
end program
