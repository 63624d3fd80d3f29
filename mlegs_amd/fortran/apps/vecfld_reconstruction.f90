!> Vector-field reconstruction tutorial (src/apps/vecfld_reconstruction.f90):
!> a q-vortex (q = 1) centred at (x, y) = (1, 1) in toroidal-poloidal form,
!> psi and chi saved, (u_r, u_phi, u_z) = tp2vec(psi, chi) saved, in
!> <flddir>{psi,chi}<ni>_FFF and {vr,vp,vz}<ni>_PPP.  Runs only at
!> NR = 200, NP = 128, NZ = 8, as the reference does.
!> Usage: vecfld_reconstruction [input_tutorial.params]
program vecfld_reconstruction
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
  if ((nr .ne. 200) .or. (np .ne. 128) .or. (nz .ne. 8)) error stop 'vecfld_reconstruction: invalid input params.'
  write(*, *) 'Initialization'
  glb_sz = [tfm%nrdim, tfm%npdim, tfm%nzdim]
  call vr%init(glb_sz, [1, 0, 2]); call vp%init(glb_sz, [1, 0, 2]); call vz%init(glb_sz, [1, 0, 2])
  call vr%set_space('PPP'); call vp%set_space('PPP'); call vz%set_space('PPP')
  call psi%init(glb_sz, [2, 1, 0]); call chi%init(glb_sz, [2, 1, 0])
  call psi%set_space('FFF'); call chi%set_space('FFF')
  call qvort_dist(psi, chi, q=1.0_p8)
  write(tag, '(i0.6)') ni
  call msave(psi, trim(adjustl(flddir)) // 'psi' // trim(tag) // '_' // psi%get_space(), is_binary=.false., is_global=.true.)
  call msave(chi, trim(adjustl(flddir)) // 'chi' // trim(tag) // '_' // chi%get_space(), is_binary=.false., is_global=.true.)
  call tp2vec(psi, chi, vr, vp, vz, tfm)
  call msave(vr, trim(adjustl(flddir)) // 'vr' // trim(tag) // '_' // vr%get_space(), is_binary=.false., is_global=.true.)
  call msave(vp, trim(adjustl(flddir)) // 'vp' // trim(tag) // '_' // vp%get_space(), is_binary=.false., is_global=.true.)
  call msave(vz, trim(adjustl(flddir)) // 'vz' // trim(tag) // '_' // vz%get_space(), is_binary=.false., is_global=.true.)
  call psi%dealloc(); call chi%dealloc()
  call vr%dealloc(); call vp%dealloc(); call vz%dealloc()
  call tfm%dealloc()
  write(*, *) 'Program Finished'

contains

  !> psi = idelsqp(-2 e^{-d^2} / (1 - x)^2), chi = idelsqp(-e^{-d^2} / q / (1 - x)^2),
  !> d the distance from (1, 1) (vecfld_reconstruction.f90, qvort_dist)
  subroutine qvort_dist(psi, chi, q)
    type(scalar), intent(inout) :: psi, chi
    real(p8), intent(in) :: q
    complex(p8), dimension(:,:,:), allocatable :: glb
    integer(i4) :: i, j, k
    real(p8) :: rr, ri
    real(p8), parameter :: xo = 1.0_p8, yo = 1.0_p8
    allocate(glb(psi%glb_sz(1), psi%glb_sz(2), psi%glb_sz(3)))
    call trans(psi, 'PPP', tfm)
    call trans(chi, 'PPP', tfm)
    glb = 0.0_p8
    do k = 1, nz
      do j = 1, np/2
        do i = 1, nr
          rr = sqrt((tfm%r(i)*cos(tfm%p(2*j-1)) - xo)**2 + (tfm%r(i)*sin(tfm%p(2*j-1)) - yo)**2)
          ri = sqrt((tfm%r(i)*cos(tfm%p(2*j  )) - xo)**2 + (tfm%r(i)*sin(tfm%p(2*j  )) - yo)**2)
          glb(i,j,k) = cmplx(-exp(-rr**2)*2.0_p8/(1.0_p8 - tfm%x(i))**2, -exp(-ri**2)*2.0_p8/(1.0_p8 - tfm%x(i))**2, p8)
        enddo
      enddo
    enddo
    call psi%disassemble(glb)
    call trans(psi, 'FFF', tfm)
    call idelsqp(psi, tfm)
    call zeroat1(psi, tfm)
    glb = 0.0_p8
    do k = 1, nz
      do j = 1, np/2
        do i = 1, nr
          rr = sqrt((tfm%r(i)*cos(tfm%p(2*j-1)) - xo)**2 + (tfm%r(i)*sin(tfm%p(2*j-1)) - yo)**2)
          ri = sqrt((tfm%r(i)*cos(tfm%p(2*j  )) - xo)**2 + (tfm%r(i)*sin(tfm%p(2*j  )) - yo)**2)
          glb(i,j,k) = cmplx(-exp(-rr**2)/q/(1.0_p8 - tfm%x(i))**2, -exp(-ri**2)/q/(1.0_p8 - tfm%x(i))**2, p8)
        enddo
      enddo
    enddo
    call chi%disassemble(glb)
    call trans(chi, 'FFF', tfm)
    call idelsqp(chi, tfm)
    call zeroat1(chi, tfm)
    deallocate(glb)
  end subroutine

end program
