!> Toroidal-poloidal projection tutorial (src/apps/tp_project.f90): loads
!> <flddir>{vr,vp,vz}<ni>_PPP (written by vecfld_reconstruction), projects
!> them with vec2tp and writes <flddir>pjpsi<ni>_FFF and pjchi<ni>_FFF.
!> Runs only at NR = 200, NP = 128, NZ = 8, as the reference does.
!> Usage: tp_project [input_tutorial.params]
program tp_project
  use mlegs_amd
  implicit none
  integer(i4) :: glb_sz(3)
  character(len=256) :: input_params_file = './input_tutorial.params'
  character(len=16) :: tag
  logical :: exists
  type(scalar) :: psi, chi, vr, vp, vz
  if (command_argument_count() .ge. 1) call get_command_argument(1, input_params_file)
  write(*, *) 'Program Started'
  inquire(file=input_params_file, exist=exists)
  if (exists) call read_input(input_params_file)
  call tfm%init()
  call colloc_pts(tfm)
  if ((nr .ne. 200) .or. (np .ne. 128) .or. (nz .ne. 8)) error stop 'tp_project: invalid input params.'
  write(*, *) 'Initialization'
  glb_sz = [tfm%nrdim, tfm%npdim, tfm%nzdim]
  call vr%init(glb_sz, [1, 0, 2]); call vp%init(glb_sz, [1, 0, 2]); call vz%init(glb_sz, [1, 0, 2])
  call vr%set_space('PPP'); call vp%set_space('PPP'); call vz%set_space('PPP')
  call psi%init(glb_sz, [2, 1, 0]); call chi%init(glb_sz, [2, 1, 0])
  call psi%set_space('FFF'); call chi%set_space('FFF')
  write(tag, '(i0.6)') ni
  call mload(trim(adjustl(flddir)) // 'vr' // trim(tag) // '_PPP', vr, is_binary=.false., is_global=.true.)
  call mload(trim(adjustl(flddir)) // 'vp' // trim(tag) // '_PPP', vp, is_binary=.false., is_global=.true.)
  call mload(trim(adjustl(flddir)) // 'vz' // trim(tag) // '_PPP', vz, is_binary=.false., is_global=.true.)
  call vec2tp(vr, vp, vz, psi, chi, tfm)
  call msave(psi, trim(adjustl(flddir)) // 'pjpsi' // trim(tag) // '_' // psi%get_space(), is_binary=.false., is_global=.true.)
  call msave(chi, trim(adjustl(flddir)) // 'pjchi' // trim(tag) // '_' // chi%get_space(), is_binary=.false., is_global=.true.)
  call psi%dealloc(); call chi%dealloc()
  call vr%dealloc(); call vp%dealloc(); call vz%dealloc()
  call tfm%dealloc()
  write(*, *) 'Program Finished'
end program
