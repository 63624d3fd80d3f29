!> 2D scalar diffusion (src/apps/scalar_diffusion_2d.f90): s(x, y, 0) =
!> max(1 - d1^8, 0.5 (1 - d2^8), 0), d1 and d2 the distances from (0.5, 0.5)
!> and (-0.5, -0.5), diffusing under visc del2 with FE-BE (no nonlinear
!> term).  Fields go to <flddir>sfld<n>_PPP every fldsavintvl steps, the
!> value at the origin to <logdir>sval_origin.dat every logsavintvl steps.
!> Usage: scalar_diffusion_2d [input_2d.params]
program scalar_diffusion_2d
  use mlegs_amd
  implicit none
  integer(i4), parameter :: stepping_notice = 1000
  integer(i4) :: glb_sz(3), fo
  character(len=256) :: input_params_file = './input_2d.params'
  logical :: exists
  complex(p8), dimension(:), allocatable :: calc
  type(scalar) :: s, nls

  if (command_argument_count() .ge. 1) call get_command_argument(1, input_params_file)
  write(*, *) 'Program Started'
  inquire(file=input_params_file, exist=exists)
  if (exists) call read_input(input_params_file)
  call timestep_set()
  call tfm%init()
  call colloc_pts(tfm)
  write(*, *) 'Initialization'

  glb_sz = [tfm%nrdim, tfm%npdim, tfm%nzdim]
  call s%init(glb_sz, [1, 0, 2]); call nls%init(glb_sz, [1, 0, 2])
  call s%set_space('PPP'); call nls%set_space('PPP')
  call init_dist(s)
  call save_field(s)
  write(*, *) 'Scalar Loaded with its Initial Distribution'
  open(newunit=fo, file=trim(adjustl(logdir)) // 'sval_origin.dat', status='replace', action='write')

  do while (curr_n .lt. ni + totaln)
    curr_n = curr_n + 1
    if (curr_n - ni .eq. totaln) dt = totaltime - (totaln - 1)*dt
    curr_t = curr_t + dt
    call nonlinear_rhs(s, nls)
    call trans(s, 'FFF', tfm); call trans(nls, 'FFF', tfm)
    call febe(s, nls, dt, tfm)
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
  call s%dealloc(); call nls%dealloc()
  call tfm%dealloc()
  write(*, *) 'Program Finished'

contains

  !> two plateaus max(1 - d^8, 0) at (0.5, 0.5) and half as high at
  !> (-0.5, -0.5), the larger of the two (scalar_diffusion_2d.f90, init_dist)
  subroutine init_dist(s)
    type(scalar), intent(inout) :: s
    complex(p8), dimension(:,:,:), allocatable :: glb
    integer(i4) :: i, j, k
    real(p8) :: rr, ri, xo, yo
    if (s%get_space() .ne. 'PPP') error stop 'init_dist: the scalar must be in PPP'
    allocate(glb(s%glb_sz(1), s%glb_sz(2), s%glb_sz(3)))
    glb = 0.0_p8
    xo = 0.5_p8; yo = 0.5_p8
    do k = 1, nz
      do j = 1, np/2
        do i = 1, nr
          rr = sqrt((tfm%r(i)*cos(tfm%p(2*j-1)) - xo)**2 + (tfm%r(i)*sin(tfm%p(2*j-1)) - yo)**2)
          ri = sqrt((tfm%r(i)*cos(tfm%p(2*j  )) - xo)**2 + (tfm%r(i)*sin(tfm%p(2*j  )) - yo)**2)
          glb(i,j,k) = cmplx(max(1.0_p8 - rr**8, 0.0_p8), max(1.0_p8 - ri**8, 0.0_p8), p8)
        enddo
      enddo
    enddo
    xo = -0.5_p8; yo = -0.5_p8
    do k = 1, nz
      do j = 1, np/2
        do i = 1, nr
          rr = sqrt((tfm%r(i)*cos(tfm%p(2*j-1)) - xo)**2 + (tfm%r(i)*sin(tfm%p(2*j-1)) - yo)**2)
          ri = sqrt((tfm%r(i)*cos(tfm%p(2*j  )) - xo)**2 + (tfm%r(i)*sin(tfm%p(2*j  )) - yo)**2)
          glb(i,j,k) = cmplx(max(real(glb(i,j,k), p8), 0.5_p8*(1.0_p8 - rr**8), 0.0_p8), &
                             max(aimag(glb(i,j,k)), 0.5_p8*(1.0_p8 - ri**8), 0.0_p8), p8)
        enddo
      enddo
    enddo
    call s%disassemble(glb)
    deallocate(glb)
  end subroutine

  !> the nonlinear term is zero: nls takes s (space and log term) with zero data
  subroutine nonlinear_rhs(s, nls)
    type(scalar), intent(in) :: s
    type(scalar), intent(inout) :: nls
    nls = s
    call nls%zero_e()
  end subroutine

  subroutine save_field(s)
    type(scalar), intent(in) :: s
    character(len=16) :: tag
    write(tag, '(i0.6)') curr_n
    call msave(s, trim(adjustl(flddir)) // 'sfld' // trim(tag) // '_' // s%get_space(), &
               is_binary=.false., is_global=.true.)
  end subroutine

end program
