!> 2D scalar transport (src/apps/scalar_transport_2d.f90): s starts at zero
!> and is carried by the azimuthal flow of a Lamb-Oseen vortex while a
!> source plateau max(1 - (4 d)^8, 0) / 2 (1 - cos(pi t)) at (0.5, 0.5)
!> feeds it, with visc del2 diffusion, FE-BE.  Fields go to
!> <flddir>sfld<n>_PPP, the value at the origin to <logdir>sval_origin.dat.
!> The nonlinear term is formed as the reference forms it: d/dphi in the
!> spectral space (column m times i m), the radial factor
!> -(1 - exp(-r)^2) / r^2 in physical space, then nls = -nls + source; the
!> per-column and per-row products run on the host (download, multiply,
!> upload), the transforms and the step on the GPU.
!> Usage: scalar_transport_2d [input_2d.params]
program scalar_transport_2d
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
  call s%zero_e()
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

  !> nls = u_phi (1/r) ds/dphi + source, with u_phi / r written as the
  !> reference writes it, (1 - exp(-r)^2) / r^2 (scalar_transport_2d.f90,
  !> nonlinear_rhs)
  subroutine nonlinear_rhs(s, nls)
    type(scalar), intent(in) :: s
    type(scalar), intent(inout) :: nls
    complex(p8), dimension(:,:,:), allocatable :: glb
    integer(i4) :: i, j, k, mm
    real(p8) :: rr, ri
    real(p8), parameter :: xo = 0.5_p8, yo = 0.5_p8
    nls = s
    call trans(nls, 'FFF', tfm)
    glb = nls%assemble()
    do mm = 1, min(size(glb, 2), tfm%chopp)
      glb(:, mm, :) = glb(:, mm, :) * cmplx(0.0_p8, real(mm - 1, p8), p8)   ! i m, m = mm - 1
    enddo
    call nls%disassemble(glb)
    call trans(nls, 'PPP', tfm)
    glb = nls%assemble()
    do k = 1, size(glb, 3)
      do j = 1, size(glb, 2)
        glb(:nr, j, k) = -glb(:nr, j, k) * (1.0_p8 - exp(-tfm%r)**2) / tfm%r**2
      enddo
    enddo
    glb = -glb
    do k = 1, nz
      do j = 1, np/2
        do i = 1, nr
          rr = sqrt((tfm%r(i)*cos(tfm%p(2*j-1)) - xo)**2 + (tfm%r(i)*sin(tfm%p(2*j-1)) - yo)**2)
          ri = sqrt((tfm%r(i)*cos(tfm%p(2*j  )) - xo)**2 + (tfm%r(i)*sin(tfm%p(2*j  )) - yo)**2)
          glb(i,j,k) = glb(i,j,k) + cmplx(max(1.0_p8 - (4*rr)**8, 0.0_p8), max(1.0_p8 - (4*ri)**8, 0.0_p8), p8) &
                       / 2.0_p8 * (1.0_p8 - cos(pi*curr_t))
        enddo
      enddo
    enddo
    call nls%disassemble(glb)
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
