!> First-order time integration tutorial (src/apps/time_integration_first.f90):
!> the 2D Lamb-Oseen vortex w(r, 0) = exp(-r^2/2) / (2 pi) diffusing under
!> visc del2 (the nonlinear term is zero) with FE-BE, run three times: with
!> the input's dt (fields saved to <flddir>sfld<n>_PPP), with dt = 0.4 and
!> with dt = 1.0 up to t = 10 (both restarted from the saved initial
!> field).  With logging on, w(0, t) is written per step to
!> <logdir>sval_origin_dt{2e-1,4e-1,1e-0}.dat; the exact value is
!> 1 / (2 pi (1 + 2 visc t)).
!> Usage: time_integration_first [input_tutorial.params]
program time_integration_first
  use mlegs_amd
  implicit none
  integer(i4), parameter :: stepping_notice = 10
  integer(i4) :: glb_sz(3), fu(3), irun
  character(len=256) :: input_params_file = './input_tutorial.params'
  character(len=22), parameter :: logname(3) = [character(len=22) :: &
    'sval_origin_dt2e-1.dat', 'sval_origin_dt4e-1.dat', 'sval_origin_dt1e-0.dat']
  real(p8), parameter :: rerun_dt(2:3) = [4.0e-1_p8, 1.0_p8]
  integer(i4), parameter :: rerun_n(2:3) = [25, 10]
  logical :: exists
  type(scalar) :: w, nlw

  if (command_argument_count() .ge. 1) call get_command_argument(1, input_params_file)
  write(*, *) 'Program Started'
  inquire(file=input_params_file, exist=exists)
  if (.not. exists) error stop 'time_integration_first: This tutorial requires the preset input params.'
  call read_input(input_params_file)
  call timestep_set()
  call tfm%init()
  call colloc_pts(tfm)
  write(*, *) 'Initialization'

  glb_sz = [tfm%nrdim, tfm%npdim, tfm%nzdim]
  call w%init(glb_sz, [1, 0, 2]); call nlw%init(glb_sz, [1, 0, 2])
  call w%set_space('PPP'); call nlw%set_space('PPP')
  call init_dist(w)
  call save_field(w)
  write(*, *) 'Scalar Loaded with its Initial Distribution'
  write(*, *) 'w(r,p,z,t=0) = 1/(2*pi) * exp(-r^2 / 2)'
  do irun = 1, 3
    open(newunit=fu(irun), file=trim(adjustl(logdir)) // logname(irun), status='replace', action='write')
  enddo

  do irun = 1, 3
    if (irun .gt. 1) then
      write(*, '(a,f4.1)') ' Re-do the simulation with dt=', rerun_dt(irun)
      dt = rerun_dt(irun); totaltime = 1.0e1_p8; ni = 0; totaln = rerun_n(irun)
      call timestep_set()
      logsavintvl = 1
      call trans(w, 'PPP', tfm)
      call mload(field_name(curr_n), w, is_binary=.false., is_global=.true.)
      call trans(nlw, 'PPP', tfm); call nlw%zero_e()
    endif
    call march(fu(irun), irun .eq. 1)
  enddo

  do irun = 1, 3
    close(fu(irun))
  enddo
  call w%dealloc(); call nlw%dealloc()
  call tfm%dealloc()
  write(*, *) 'Program Finished'

contains

  !> FE-BE steps to ni + totaln (time_integration_first.f90, time stepping loops)
  subroutine march(fo, save)
    integer, intent(in) :: fo
    logical, intent(in) :: save
    complex(p8), dimension(:), allocatable :: calc
    do while (curr_n .lt. ni + totaln)
      curr_n = curr_n + 1
      if (curr_n - ni .eq. totaln) dt = totaltime - (totaln - 1)*dt
      curr_t = curr_t + dt
      call nonlinear_rhs(w, nlw)
      call trans(w, 'FFF', tfm); call trans(nlw, 'FFF', tfm)
      call febe(w, nlw, dt, tfm)
      if (save .and. isfldsav) then
        call trans(w, 'PPP', tfm)
        if (mod(curr_n, fldsavintvl) .eq. 0) call save_field(w)
      endif
      if (islogsav) then
        if (mod(curr_n, logsavintvl) .eq. 0) then
          calc = calcat0(w, tfm)
          write(fo, '(i10, " ", f10.6, " ", 1pe23.14)') curr_n, curr_t, real(calc(1), p8)
        endif
      endif
      if (mod(curr_n, stepping_notice) .eq. 0) &
        write(*, '(a,i10,a,f11.6,a)') ' time stepping # ', curr_n, ' (t: ', curr_t, ')'
    enddo
  end subroutine

  !> w = exp(-r^2/2) / (2 pi) at every collocation point (both packed angles)
  subroutine init_dist(s)
    type(scalar), intent(inout) :: s
    complex(p8), dimension(:,:,:), allocatable :: glb
    integer(i4) :: i, j, k
    real(p8) :: rr, ri
    if (s%get_space() .ne. 'PPP') error stop 'init_dist: the scalar must be in PPP'
    allocate(glb(s%glb_sz(1), s%glb_sz(2), s%glb_sz(3)))
    glb = 0.0_p8
    do k = 1, nz
      do j = 1, np/2
        do i = 1, nr
          rr = sqrt((tfm%r(i)*cos(tfm%p(2*j-1)))**2 + (tfm%r(i)*sin(tfm%p(2*j-1)))**2)
          ri = sqrt((tfm%r(i)*cos(tfm%p(2*j  )))**2 + (tfm%r(i)*sin(tfm%p(2*j  )))**2)
          glb(i,j,k) = cmplx(exp(-rr**2/2.0_p8), exp(-ri**2/2.0_p8), p8) / (2.0_p8*pi)
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

  function field_name(n) result(fn)
    integer(i4), intent(in) :: n
    character(len=:), allocatable :: fn
    character(len=16) :: tag
    write(tag, '(i0.6)') n
    fn = trim(adjustl(flddir)) // 'sfld' // trim(tag) // '_PPP'
  end function

  subroutine save_field(s)
    type(scalar), intent(in) :: s
    call msave(s, field_name(curr_n), is_binary=.false., is_global=.true.)
  end subroutine

end program
