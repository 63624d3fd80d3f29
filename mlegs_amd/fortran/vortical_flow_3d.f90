!> Temporal evolution of a 3D vortex pair in a uniform axial flow, one MI355X.
!>
!> The program of the reference's src/apps/vortical_flow_3d.f90 -- same
!> input.params, initial condition (two q-vortices at x = -2 and +2,
!> q = 1, uniform U = b e_z with b = -0.5), FE-BE + Richardson bootstrap and
!> AB2/CN loop with de-aliasing, SVV and zeroat1 -- written against the
!> mlegs_amd module, so every field operation runs on the GPU.
!> Differences: one process drives one GPU (no MPI); the 1e-10 noise of the
!> initial condition is the library's seeded spectral perturbation instead of
!> the compiler's rand(), so runs are reproducible.  With ISFLDSAV the
!> vorticity-magnitude snapshots vormagfld<n>_PPP are written to FLDDIR, as
!> the reference's save_vort_mag writes them.
!>
!> Usage:  vortical_flow_3d [input.params [dump_prefix [precision]]]
!>   input.params  default ./input.params (the built-in defaults when absent)
!>   dump_prefix   write the final psi and chi to <prefix>_psi.bin and
!>                 <prefix>_chi.bin: three int64 sizes, one float64 ln, then
!>                 the global complex(8) array e(nrdim, npdim, nzdim)
!>   precision     Legendre contraction precision: 0 FP64 (default),
!>                 1 FP32 MFMA, 2 split-BF16 MFMA, 3 FP64 MFMA, or 'mixed':
!>                 FP32 MFMA with the projection's column m = 0 on FP64 MFMA
program vortical_flow_3d
  use iso_c_binding
  use mlegs_amd
  implicit none

  interface
    integer(c_int) function mlx_seed_spectral(host, n1, n2, n3, chops, seed, field_id, amp, scale, kind) &
        bind(C, name='mlx_seed_spectral')
      import :: c_int, c_double_complex, c_int64_t, c_int32_t, c_double
      complex(c_double_complex), intent(out) :: host(*)
      integer(c_int64_t), value :: n1, n2, n3
      integer(c_int32_t), intent(in) :: chops(5)
      integer(c_int64_t), value :: seed
      integer(c_int32_t), value :: field_id
      real(c_double), value :: amp, scale
      integer(c_int32_t), value :: kind
    end function
  end interface

  integer(i4), parameter :: stepping_notice = 1
  integer(i8), parameter :: seed = 20261202_i8
  real(p8) :: dt_step, svv_gain_psi = 0.0_p8, svv_gain_chi = 0.0_p8, ln_keep
  integer(i4) :: glb_sz(3), i, nargs, ios
  integer(i8) :: c0, c1, crate
  character(len=256) :: input_params_file = './input.params', dump_prefix = '', arg
  logical :: exists, mixed = .false.
  type(scalar) :: uz
  type(scalar) :: psi, nlpsi, chi, nlchi, psi_prev, nlpsi_prev, chi_prev, nlchi_prev
  type(scalar) :: psi_rich, chi_rich

  nargs = command_argument_count()
  if (nargs .ge. 1) call get_command_argument(1, input_params_file)
  if (nargs .ge. 2) call get_command_argument(2, dump_prefix)
  if (nargs .ge. 3) then
    call get_command_argument(3, arg)
    if (trim(arg) .eq. 'mixed') then
      mixed = .true.
      mlx_precision = MLX_PREC_FP32
    else
      read(arg, *, iostat=ios) mlx_precision
      if (ios .ne. 0) error stop "vortical_flow_3d: precision must be an integer or 'mixed'"
    endif
  endif
  write(*, *) 'Program Started'
  call system_clock(c0, crate)

  inquire(file=input_params_file, exist=exists)
  if (exists) call read_input(input_params_file)
  call timestep_set()
  call tfm%init()
  if (mixed) call tfm%set_stage_precision(int(MLX_STAGE_PROJ0, i4), int(MLX_PREC_FP64_MFMA, i4))
  write(*, '(a,3i6,a,i2,a,l2)') ' Initialization: NR NP NZ =', nr, np, nz, ', precision', mlx_precision, &
                                ', mixed', mixed

  glb_sz = [tfm%nrdim, tfm%npdim, tfm%nzdim]
  call uz%init(glb_sz, [1, 0, 2]); call uz%set_space('PPP')
  call psi%init(glb_sz, [2, 1, 0]); call nlpsi%init(glb_sz, [2, 1, 0])
  call psi%set_space('FFF'); call nlpsi%set_space('FFF')
  call psi_prev%init(glb_sz, [2, 1, 0]); call nlpsi_prev%init(glb_sz, [2, 1, 0])
  call psi_prev%set_space('FFF'); call nlpsi_prev%set_space('FFF')
  call chi%init(glb_sz, [2, 1, 0]); call nlchi%init(glb_sz, [2, 1, 0])
  call chi%set_space('FFF'); call nlchi%set_space('FFF')
  call chi_prev%init(glb_sz, [2, 1, 0]); call nlchi_prev%init(glb_sz, [2, 1, 0])
  call chi_prev%set_space('FFF'); call nlchi_prev%set_space('FFF')
  call psi_rich%init(glb_sz, [2, 1, 0]); call psi_rich%set_space('FFF')
  call chi_rich%init(glb_sz, [2, 1, 0]); call chi_rich%set_space('FFF')

  call qvort_dist_tp(psi, chi, q=1.0_p8, ran_noise=1.0e-10_p8)
  call uniform_z_fld(uz, b=-0.5_p8)
  write(*, *) 'Swirling Flow Field Declaration & Initialization'
  if (isfldsav) call save_vort_mag(psi, chi)

  !> Richardson extrapolation bootstrapping (vortical_flow_3d.f90:116-147)
  call advection_rhs(psi, chi, nlpsi, nlchi, uz)
  psi_prev = psi; nlpsi_prev = nlpsi
  chi_prev = chi; nlchi_prev = nlchi
  psi_rich = psi
  chi_rich = chi
  do i = 1, 1
    call advection_rhs(psi_rich, chi_rich, nlpsi, nlchi, uz)
    call febe(psi_rich, nlpsi, dt, tfm)
    call febe(chi_rich, nlchi, dt, tfm)
  enddo
  do i = 1, 2
    call advection_rhs(psi, chi, nlpsi, nlchi, uz)
    call febe(psi, nlpsi, dt/2.0_p8, tfm)
    call febe(chi, nlchi, dt/2.0_p8, tfm)
  enddo
  ln_keep = psi%get_ln()
  call psi%axpby(2.0_p8/1.0_p8, psi_rich, -1.0_p8/1.0_p8)   ! psi%e = 2 psi%e - psi_rich%e
  call psi%set_ln(ln_keep)
  ln_keep = chi%get_ln()
  call chi%axpby(2.0_p8/1.0_p8, chi_rich, -1.0_p8/1.0_p8)
  call chi%set_ln(ln_keep)
  call dealias(psi, tfm); call dealias(chi, tfm)
  call svv_filter(psi, tfm, svv_gain_psi); call svv_filter(chi, tfm, svv_gain_chi)
  call zeroat1(psi, tfm); call zeroat1(chi, tfm)
  call advection_rhs(psi, chi, nlpsi, nlchi, uz)
  call psi_rich%dealloc()
  call chi_rich%dealloc()
  curr_n = 1; curr_t = dt
  call tfm%sync()
  write(*, *) 'Richardson Extrapolation (time stepping # 1)'
  if (isfldsav) call save_vort_mag(psi, chi)

  !> main time stepping (vortical_flow_3d.f90:160-206)
  call system_clock(c0)
  do while (curr_n .lt. ni + totaln)
    curr_n = curr_n + 1
    dt_step = dt
    if (curr_n - ni .eq. totaln) dt_step = totaltime - (totaln - 1)*dt
    curr_t = curr_t + dt_step
    call trans(psi, 'FFF', tfm); call trans(chi, 'FFF', tfm)
    call trans(nlpsi, 'FFF', tfm); call trans(nlchi, 'FFF', tfm)
    call trans(psi_prev, 'FFF', tfm); call trans(chi_prev, 'FFF', tfm)
    call trans(nlpsi_prev, 'FFF', tfm); call trans(nlchi_prev, 'FFF', tfm)
    call abcn(psi, psi_prev, nlpsi, nlpsi_prev, dt_step, tfm)
    call abcn(chi, chi_prev, nlchi, nlchi_prev, dt_step, tfm)
    call dealias(psi, tfm); call dealias(chi, tfm)
    call svv_filter(psi, tfm, svv_gain_psi); call svv_filter(chi, tfm, svv_gain_chi)
    call zeroat1(psi, tfm); call zeroat1(chi, tfm)
    call advection_rhs(psi, chi, nlpsi, nlchi, uz)
    call check_stability(psi, chi)
    if (isfldsav) then
      if (mod(curr_n, fldsavintvl) .eq. 0) call save_vort_mag(psi, chi)
    endif
    if (mod(curr_n, stepping_notice) .eq. 0) then
      call tfm%sync()
      call system_clock(c1)
      write(*, '(a,i10,a,f11.6,a,f13.6,a)') ' time stepping # ', curr_n, ' (t: ', curr_t, ') ', &
        real(c1 - c0, p8) / real(crate, p8) / stepping_notice, ' seconds per step'
      c0 = c1
    endif
  enddo

  if (len_trim(dump_prefix) .gt. 0) then
    call dump(psi, trim(dump_prefix) // '_psi.bin')
    call dump(chi, trim(dump_prefix) // '_chi.bin')
  endif
  write(*, '(a,2es23.14)') ' svv gains: ', svv_gain_psi, svv_gain_chi

  call uz%dealloc()
  call psi%dealloc(); call nlpsi%dealloc()
  call psi_prev%dealloc(); call nlpsi_prev%dealloc()
  call chi%dealloc(); call nlchi%dealloc()
  call chi_prev%dealloc(); call nlchi_prev%dealloc()
  call tfm%dealloc()
  write(*, *) 'Program Finished'

contains

  !> two q-vortices at (x, y) = (-2, 0) and (2, 0) (vortical_flow_3d.f90:258-326)
  subroutine qvort_dist_tp(psi, chi, q, ran_noise)
    type(scalar), intent(inout) :: psi, chi
    real(p8), intent(in) :: q, ran_noise
    complex(p8), dimension(:,:,:), allocatable :: glb
    integer(i4) :: i, j, k, xo, yo
    real(p8) :: rr, ri
    if (psi%get_space() .ne. 'FFF') error stop 'qvort_dist_tp: psi must be in FFF'
    if (chi%get_space() .ne. 'FFF') error stop 'qvort_dist_tp: chi must be in FFF'
    allocate(glb(psi%glb_sz(1), psi%glb_sz(2), psi%glb_sz(3)))
    call trans(psi, 'PPP', tfm)
    call trans(chi, 'PPP', tfm)
    glb = 0.0_p8
    do k = 1, nz
      do j = 1, np/2
        do i = 1, nr
          do xo = -2, 2, 4
            yo = 0
            rr = sqrt((tfm%r(i)*cos(tfm%p(2*j-1)) - xo)**2 + (tfm%r(i)*sin(tfm%p(2*j-1)) - yo)**2)
            ri = sqrt((tfm%r(i)*cos(tfm%p(2*j  )) - xo)**2 + (tfm%r(i)*sin(tfm%p(2*j  )) - yo)**2)
            glb(i,j,k) = glb(i,j,k) + cmplx(-exp(-rr**2)*2.0_p8/(1.0_p8 - tfm%x(i))**2, &
                                            -exp(-ri**2)*2.0_p8/(1.0_p8 - tfm%x(i))**2, p8)
          enddo
        enddo
      enddo
    enddo
    call psi%disassemble(glb)
    call trans(psi, 'FFF', tfm)
    call idelsqp(psi, tfm)
    call add_noise(psi, ran_noise, 1)
    call zeroat1(psi, tfm)
    glb = 0.0_p8
    do k = 1, nz
      do j = 1, np/2
        do i = 1, nr
          do xo = -2, 2, 4
            yo = 0
            rr = sqrt((tfm%r(i)*cos(tfm%p(2*j-1)) - xo)**2 + (tfm%r(i)*sin(tfm%p(2*j-1)) - yo)**2)
            ri = sqrt((tfm%r(i)*cos(tfm%p(2*j  )) - xo)**2 + (tfm%r(i)*sin(tfm%p(2*j  )) - yo)**2)
            glb(i,j,k) = glb(i,j,k) + cmplx(-exp(-rr**2)/q/(1.0_p8 - tfm%x(i))**2, &
                                            -exp(-ri**2)/q/(1.0_p8 - tfm%x(i))**2, p8)
          enddo
        enddo
      enddo
    enddo
    call chi%disassemble(glb)
    call trans(chi, 'FFF', tfm)
    call idelsqp(chi, tfm)
    call add_noise(chi, ran_noise, 2)
    call zeroat1(chi, tfm)
    deallocate(glb)
  end subroutine

  !> s = s + seeded admissible perturbation of amplitude amp (FFF)
  subroutine add_noise(s, amp, field_id)
    type(scalar), intent(inout) :: s
    real(p8), intent(in) :: amp
    integer(i4), intent(in) :: field_id
    type(scalar) :: pert
    complex(p8), dimension(:,:,:), allocatable :: glb
    integer(c_int32_t) :: chops(5)
    if (amp .eq. 0.0_p8) return
    allocate(glb(s%glb_sz(1), s%glb_sz(2), s%glb_sz(3)))
    chops = [nrchop, npchop, nzchop, np, nz]
    if (mlx_seed_spectral(glb, int(s%glb_sz(1), c_int64_t), int(s%glb_sz(2), c_int64_t), &
                          int(s%glb_sz(3), c_int64_t), chops, seed, int(field_id, c_int32_t), amp, 4.0_p8, 0) .ne. 0) &
      error stop 'add_noise: seeding failed'
    call pert%init(s%glb_sz, [2, 1, 0])
    call pert%set_space('FFF')
    call pert%disassemble(glb)
    call s%axpby(1.0_p8, pert, 1.0_p8)
    call pert%dealloc()
    deallocate(glb)
  end subroutine

  !> uniform axial flow U = b e_z (vortical_flow_3d.f90:328-351)
  subroutine uniform_z_fld(uz, b)
    type(scalar), intent(inout) :: uz
    real(p8) :: b
    complex(p8), dimension(:,:,:), allocatable :: glb
    if (uz%get_space() .ne. 'PPP') error stop 'uniform_z_fld: uz must be in PPP'
    allocate(glb(uz%glb_sz(1), uz%glb_sz(2), uz%glb_sz(3)))
    glb = 0.0_p8
    glb(1:nr, 1:np/2, 1:nz) = cmplx(b, b, p8)
    call uz%disassemble(glb)
    deallocate(glb)
  end subroutine

  !> the vorticity magnitude field at time step curr_n, written as
  !> <flddir>vormagfld<curr_n, 6 digits>_PPP in the msave ASCII format
  !> (vortical_flow_3d.f90:411-459)
  subroutine save_vort_mag(psi, chi)
    type(scalar), intent(inout) :: psi, chi
    type(scalar) :: vormag
    character(len=16) :: tag
    if (psi%get_space() .ne. 'FFF') error stop 'save_vort_mag: psi must be in FFF'
    if (chi%get_space() .ne. 'FFF') error stop 'save_vort_mag: chi must be in FFF'
    call vormag%init(glb_sz, [1, 0, 2])
    call vormag%set_space('PPP')
    call vorticity_magnitude(psi, chi, vormag)
    write(tag, '(i0.6)') curr_n
    call msave(vormag, trim(adjustl(flddir)) // 'vormagfld' // trim(tag) // '_' // vormag%get_space(), &
               is_binary=.false., is_global=.true.)
    write(*, *) 'Field Data Saved at Time Step #' // trim(tag)
    call vormag%dealloc()
  end subroutine

  subroutine dump(s, fn)
    type(scalar), intent(in) :: s
    character(len=*), intent(in) :: fn
    complex(p8), dimension(:,:,:), allocatable :: glb
    integer :: fu
    glb = s%assemble()
    open(newunit=fu, file=fn, access='stream', form='unformatted', status='replace', action='write')
    write(fu) int(shape(glb), i8), s%get_ln(), glb
    close(fu)
  end subroutine

end program vortical_flow_3d
