!> 1D radial wave propagation (src/apps/wave_propagation_1d.f90):
!> d^2 s/dt^2 = del2 s with s(r, 0) = exp(-(r/4)^2), ds/dt(r, 0) = 0, as the
!> first-order system s' = dsdt, dsdt' = del2 s, both advanced by fefe on
!> the radial kit (NP = NZ = 1).  Fields go to <flddir>sfld<n>_PPP, the
!> value at the origin to <logdir>sval_origin.dat.
!> Usage: wave_propagation_1d [input_1d.params]
program wave_propagation_1d
  use mlegs_amd
  implicit none
  integer(i4), parameter :: stepping_notice = 100000
  integer(i4) :: glb_sz(3), fo
  character(len=256) :: input_params_file = './input_1d.params'
  logical :: exists
  complex(p8), dimension(:), allocatable :: calc
  type(scalar) :: s, dsdt, del2s

  if (command_argument_count() .ge. 1) call get_command_argument(1, input_params_file)
  write(*, *) 'Program Started'
  inquire(file=input_params_file, exist=exists)
  if (exists) call read_input(input_params_file)
  call timestep_set()
  call tfm%init()
  call colloc_pts(tfm)
  write(*, *) 'Initialization'

  glb_sz = [tfm%nrdim, tfm%npdim, tfm%nzdim]
  call s%init(glb_sz, [0, 1, 2]); call dsdt%init(glb_sz, [0, 1, 2]); call del2s%init(glb_sz, [0, 1, 2])
  call s%set_space('PPP'); call dsdt%set_space('PPP'); call del2s%set_space('PPP')
  call init_dist(s, dsdt)
  call save_field(s)
  write(*, *) 'Scalar Loaded with its Initial Distribution'
  write(*, *) 's(r,t=0) = exp(-(r/4)^2)'
  open(newunit=fo, file=trim(adjustl(logdir)) // 'sval_origin.dat', status='replace', action='write')

  do while (curr_n .lt. ni + totaln)
    curr_n = curr_n + 1
    if (curr_n - ni .eq. totaln) dt = totaltime - (totaln - 1)*dt
    curr_t = curr_t + dt
    call trans(s, 'FFF', tfm)
    del2s = s
    call del2(del2s, tfm)
    call trans(s, 'FFF', tfm); call trans(dsdt, 'FFF', tfm); call trans(del2s, 'FFF', tfm)
    call fefe(s, dsdt, dt, tfm)
    call fefe(dsdt, del2s, dt, tfm)
    if (isfldsav) then
      call trans(s, 'PPP', tfm)
      if (mod(curr_n, fldsavintvl) .eq. 0) call save_field(s)
    endif
    if (islogsav) then
      if (mod(curr_n, logsavintvl) .eq. 0) then
        calc = calcat0(s, tfm)
        write(fo, '(i10, " ", f10.6, " ", 1pe13.4)') curr_n, curr_t, real(calc(1), p8)
      endif
    endif
    if (mod(curr_n, stepping_notice) .eq. 0) &
      write(*, '(a,i10,a,f11.6,a)') ' time stepping # ', curr_n, ' (t: ', curr_t, ')'
  enddo

  close(fo)
  call s%dealloc(); call dsdt%dealloc(); call del2s%dealloc()
  call tfm%dealloc()
  write(*, *) 'Program Finished'

contains

  !> s = exp(-(r/4)^2) in the single column, dsdt = 0
  subroutine init_dist(s, dsdt)
    type(scalar), intent(inout) :: s, dsdt
    complex(p8), dimension(:,:,:), allocatable :: glb
    integer(i4) :: i
    if (s%get_space() .ne. 'PPP') error stop 'init_dist: the scalar must be in PPP'
    allocate(glb(s%glb_sz(1), s%glb_sz(2), s%glb_sz(3)))
    glb = 0.0_p8
    do i = 1, nr
      glb(i, 1, 1) = exp(-((tfm%r(i)/4.0_p8)**2))
    enddo
    call s%disassemble(glb)
    call dsdt%zero_e()
    deallocate(glb)
  end subroutine

  subroutine save_field(s)
    type(scalar), intent(in) :: s
    character(len=16) :: tag
    write(tag, '(i0.6)') curr_n
    call msave(s, trim(adjustl(flddir)) // 'sfld' // trim(tag) // '_' // s%get_space(), &
               is_binary=.false., is_global=.true.)
  end subroutine

end program
