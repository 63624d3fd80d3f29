!> The skeleton every program starts from (src/apps/barebone_template.f90):
!> read input.params when present, set up the transformation kit on the GPU,
!> release it.
!> Usage: barebone_template [input.params]
program barebone_template
  use mlegs_amd
  implicit none
  character(len=256) :: input_params_file = './input.params'
  logical :: exists
  if (command_argument_count() .ge. 1) call get_command_argument(1, input_params_file)
  write(*, *) 'Program Started'
  inquire(file=input_params_file, exist=exists)
  if (exists) call read_input(input_params_file)
  call tfm%init()
  write(*, *) 'Initialization'
  call tfm%dealloc()
  write(*, *) 'Program Finished'
end program
