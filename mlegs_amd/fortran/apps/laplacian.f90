!> Laplacian tutorial (src/apps/laplacian.f90): loads <logdir>scalar_FFF.dat
!> (or, when absent, sets the FFF entries (2,2,1) and (3,3,1) to 1 and saves
!> them there), applies del2 and writes scalar_FFF_Laplacian.dat and, after
!> the backward transform, scalar_PPP_Laplacian.dat.
!> Usage: laplacian [input_2d.params]
program laplacian
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
  write(*, *) 'Initialization'
  glb_sz = [tfm%nrdim, tfm%npdim, tfm%nzdim]
  call s%init(glb_sz, [0, 1, 2])
  call s%set_space('FFF')
  inquire(file=trim(adjustl(logdir)) // 'scalar_FFF.dat', exist=exists)
  if (exists) then
    call mload(trim(adjustl(logdir)) // 'scalar_FFF.dat', s, is_binary=.false., is_global=.true.)
  else
    allocate(array_glb(glb_sz(1), glb_sz(2), glb_sz(3)))
    array_glb = 0.0_p8
    if (glb_sz(2) .ge. 2) array_glb(2, 2, 1) = 1.0_p8
    if (glb_sz(2) .ge. 3) array_glb(3, 3, 1) = 1.0_p8
    call s%disassemble(array_glb)
    deallocate(array_glb)
    call msave(s, trim(adjustl(logdir)) // 'scalar_FFF.dat', is_binary=.false., is_global=.true.)
  endif
  call del2(s, tfm)
  call msave(s, trim(adjustl(logdir)) // 'scalar_FFF_Laplacian.dat', is_binary=.false., is_global=.true.)
  call trans(s, 'PPP', tfm)
  call msave(s, trim(adjustl(logdir)) // 'scalar_PPP_Laplacian.dat', is_binary=.false., is_global=.true.)
  call s%dealloc()
  call tfm%dealloc()
  write(*, *) 'Program Finished'
end program
