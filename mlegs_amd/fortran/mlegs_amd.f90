!> Fortran host interface of the MI355X spectral-transform library.
!>
!> One module gives a Fortran program the call surface of the reference's
!> modules for the time-step path:
!>   mlegs_base    run parameters with their defaults (src/modules/mlegs_base.f90:7-53)
!>   mlegs_misc    read_input, timestep_set (src/submodules/mlegs_misc_io.f90:307-408)
!>   mlegs_spectfm type(tfm_kit_3d) and the global kit tfm (src/modules/mlegs_spectfm.f90:14-63)
!>   mlegs_scalar  type(scalar) with init/dealloc/chop_offset/assemble/disassemble,
!>                 assignment(=) and the generic operators trans, chop, dealias,
!>                 svv_filter, calcat0/1, zeroat1, delsqp, idelsqp, xxdx, del2h,
!>                 del2, idel2, ihelm, helmp, ihelmp, fefe, abab, febe, abcn, vecprod, vec2tp, tp2vec,
!>                 tp2curlvec (src/modules/mlegs_scalar.f90:15-428)
!> Every operator calls the C ABI of include/mlegs_amd.h through iso_c_binding;
!> the data of a scalar stays in GPU memory.  Differences from the reference
!> surface, all forced by the data living on the device:
!>   - s%e is not a host array: s%assemble() returns the global array and
!>     s%disassemble(a) stores one (the reference's MPI gather/scatter names);
!>   - s%space and s%ln are read with s%get_space() / s%get_ln() and set with
!>     s%set_space() / s%set_ln();
!>   - s%init(glb_sz, axis_comm) ignores axis_comm: one process drives one GPU
!>     and holds the whole field.
!> A failing call stops the program with the library's message (error stop),
!> as the reference's error stop / MPI_abort do.
module mlegs_amd
  use iso_c_binding
  use iso_fortran_env, only: error_unit
  implicit none
  private

  integer, parameter, public :: i4 = selected_int_kind(9)
  integer, parameter, public :: i8 = selected_int_kind(18)
  integer, parameter, public :: p8 = selected_real_kind(15, 307)
  real(p8), parameter, public :: pi = 3.14159265358979323846264338327950288_p8

  integer(c_int), parameter, public :: MLX_PREC_FP64 = 0, MLX_PREC_FP32 = 1, MLX_PREC_BF16X3 = 2
  integer(c_int), parameter, public :: MLX_PREC_FP64_MFMA = 3
  !> contraction sites with their own precision (kit%set_stage_precision)
  integer(c_int), parameter, public :: MLX_STAGE_SYNTH = 0, MLX_STAGE_UP = 1, MLX_STAGE_PROJ = 2, &
                                        MLX_STAGE_PROJ0 = 3

  !> run parameters (mlegs_base.f90:7-53) and their defaults
  integer(i4), public :: nr = 256, np = 128, nz = 128
  integer(i4), public :: nrchop = 256, npchop = 65, nzchop = 65
  real(p8), public :: ell = 4.0_p8, zlen = 2.0_p8 * pi
  real(p8), public :: dt = 1.0e-3_p8, ti = 0.0_p8, totaltime = 1.0e2_p8
  integer(i4), public :: ni = 0, totaln = 100000
  real(p8), public :: visc = 1.0e-3_p8
  integer(i4), public :: hyperpow = 0
  real(p8), public :: hypervisc = 0.0_p8
  logical, public :: is_svv = .true.
  real(p8), public :: svv_cutoff = 0.75_p8, svv_target = 2.0e-2_p8, svv_strength = 0.12_p8, svv_relax = 0.25_p8
  character(len=256), public :: flddir = './output/fld/', logdir = './output/log/'
  logical, public :: isfldsav = .true., islogsav = .true.
  integer(i4), public :: fldsavintvl = 100, logsavintvl = 1
  integer(i4), public :: curr_n = 0
  real(p8), public :: curr_t = 0.0_p8
  !> HIP device of the global kit and precision of its Legendre contractions
  integer(i4), public :: mlx_device = 0
  integer(i4), public :: mlx_precision = MLX_PREC_FP64

  !> mlx_params of include/mlegs_amd.h
  type, bind(C) :: mlx_params_c
    integer(c_int32_t) :: nr, np, nz, nrchop, npchop, nzchop
    real(c_double) :: ell, zlen, dt, ti, totaltime
    integer(c_int32_t) :: ni, totaln
    real(c_double) :: visc
    integer(c_int32_t) :: hyperpow
    real(c_double) :: hypervisc
    integer(c_int32_t) :: is_svv
    real(c_double) :: svv_cutoff, svv_target, svv_strength, svv_relax
  end type

  !> the 3D transformation kit (one library context on one GPU)
  type, public :: tfm_kit_3d
    type(c_ptr) :: ctx = c_null_ptr
    logical :: is_set = .false.
    integer(i4) :: nrdim = 0, npdim = 0, nzdim = 0, nrh = 0, ne = 0, chopzl = 0, chopzu = 0, nph = 0, chopp = 0
    real(p8), allocatable :: x(:), w(:), ln(:), r(:), p(:), z(:), ak(:)
  contains
    procedure :: init => kit_init
    procedure :: dealloc => kit_dealloc
    procedure :: set_precision => kit_set_precision
    procedure :: set_stage_precision => kit_set_stage_precision
    procedure :: sync => kit_sync
  end type
  type(tfm_kit_3d), public, target :: tfm

  !> a scalar field on the device
  type, public :: scalar
    type(c_ptr) :: h = c_null_ptr
    type(c_ptr) :: ctx = c_null_ptr
    integer(i4) :: glb_sz(3) = 0
  contains
    procedure :: init => scalar_init
    procedure :: dealloc => scalar_dealloc
    procedure :: chop_offset => scalar_chop_offset
    procedure :: assemble => scalar_assemble
    procedure :: disassemble => scalar_disassemble
    procedure :: get_space => scalar_get_space
    procedure :: set_space => scalar_set_space
    procedure :: get_ln => scalar_get_ln
    procedure :: set_ln => scalar_set_ln
    procedure :: axpby => scalar_axpby
    procedure :: zero_e => scalar_zero_e
  end type

  public :: read_input, timestep_set
  public :: assignment(=)
  public :: trans, chop, dealias, svv_filter, calcat0, calcat1, zeroat1, fftreat
  public :: delsqp, idelsqp, xxdx, del2h, del2, idel2, ihelm, helmp, ihelmp, fefe, abab, febe, abcn
  public :: vecprod, vec2tp, tp2vec, tp2curlvec, advection_rhs, check_stability, energy
  public :: msave, mload, vorticity_magnitude, colloc_pts

  interface assignment(=)
    module procedure scalar_copy
  end interface
  interface msave
    module procedure msave_scalar, msave_real1
  end interface
  interface mload
    module procedure mload_scalar
  end interface
  interface trans
    module procedure scalar_trans
  end interface
  interface chop
    module procedure scalar_chop
  end interface
  interface dealias
    module procedure scalar_dealias
  end interface
  interface svv_filter
    module procedure scalar_svv_filter
  end interface
  interface calcat0
    module procedure scalar_calcat0
  end interface
  interface calcat1
    module procedure scalar_calcat1
  end interface
  interface zeroat1
    module procedure scalar_zeroat1
  end interface
  interface fftreat
    module procedure scalar_fftreat
  end interface
  interface delsqp
    module procedure scalar_delsqp
  end interface
  interface idelsqp
    module procedure scalar_idelsqp
  end interface
  interface xxdx
    module procedure scalar_xxdx
  end interface
  interface del2h
    module procedure scalar_del2h
  end interface
  interface del2
    module procedure scalar_del2
  end interface
  interface idel2
    module procedure scalar_idel2_proln, scalar_idel2_preln
  end interface
  interface ihelm
    module procedure scalar_ihelm
  end interface
  interface helmp
    module procedure scalar_helmp
  end interface
  interface ihelmp
    module procedure scalar_ihelmp
  end interface
  interface fefe
    module procedure scalar_fefe
  end interface
  interface abab
    module procedure scalar_abab
  end interface
  interface febe
    module procedure scalar_febe
  end interface
  interface abcn
    module procedure scalar_abcn
  end interface
  interface vecprod
    module procedure vector_product
  end interface
  interface vec2tp
    module procedure vector_projection
  end interface
  interface tp2vec
    module procedure vector_reconstruction
  end interface
  interface tp2curlvec
    module procedure curl_vector_reconstruction
  end interface

  !> the C ABI (include/mlegs_amd.h)
  interface
    integer(c_int) function mlx_ctx_create(p, device, stream, out) bind(C, name='mlx_ctx_create')
      import :: c_int, c_ptr, mlx_params_c
      type(mlx_params_c), intent(in) :: p
      integer(c_int), value :: device
      type(c_ptr), value :: stream
      type(c_ptr), intent(out) :: out
    end function
    integer(c_int) function mlx_ctx_destroy(ctx) bind(C, name='mlx_ctx_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    type(c_ptr) function mlx_last_error(ctx) bind(C, name='mlx_last_error')
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function mlx_sync(ctx) bind(C, name='mlx_sync')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function mlx_kit_init(ctx) bind(C, name='mlx_kit_init')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function mlx_kit_dims(ctx, out9) bind(C, name='mlx_kit_dims')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(out) :: out9(9)
    end function
    integer(c_int) function mlx_kit_get(ctx, which, host, n) bind(C, name='mlx_kit_get')
      import :: c_int, c_ptr, c_char, c_double, c_int64_t
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: which(*)
      real(c_double), intent(out) :: host(*)
      integer(c_int64_t), value :: n
    end function
    integer(c_int) function mlx_set_precision(ctx, prec) bind(C, name='mlx_set_precision')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: prec
    end function
    integer(c_int) function mlx_set_stage_precision(ctx, stage, prec) bind(C, name='mlx_set_stage_precision')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: stage, prec
    end function
    integer(c_int) function mlx_field_create(ctx, out) bind(C, name='mlx_field_create')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      type(c_ptr), intent(out) :: out
    end function
    integer(c_int) function mlx_field_destroy(ctx, f) bind(C, name='mlx_field_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, f
    end function
    integer(c_int) function mlx_field_copy(ctx, dst, src) bind(C, name='mlx_field_copy')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, dst, src
    end function
    integer(c_int) function mlx_field_upload(ctx, f, host, n1, n2, n3, space, ln, offsets) &
        bind(C, name='mlx_field_upload')
      import :: c_int, c_ptr, c_double_complex, c_int64_t, c_char, c_double, c_int32_t
      type(c_ptr), value :: ctx, f
      complex(c_double_complex), intent(in) :: host(*)
      integer(c_int64_t), value :: n1, n2, n3
      character(kind=c_char), intent(in) :: space(*)
      real(c_double), value :: ln
      integer(c_int32_t), intent(in) :: offsets(3)
    end function
    integer(c_int) function mlx_field_download(ctx, f, host, n1, n2, n3, space4, ln, offsets) &
        bind(C, name='mlx_field_download')
      import :: c_int, c_ptr, c_double_complex, c_int64_t, c_char, c_double, c_int32_t
      type(c_ptr), value :: ctx, f
      complex(c_double_complex), intent(out) :: host(*)
      integer(c_int64_t), value :: n1, n2, n3
      character(kind=c_char), intent(out) :: space4(4)
      real(c_double), intent(out) :: ln
      integer(c_int32_t), intent(out) :: offsets(3)
    end function
    integer(c_int) function mlx_field_set_meta(ctx, f, space, ln, offsets) bind(C, name='mlx_field_set_meta')
      import :: c_int, c_ptr, c_char, c_double, c_int32_t
      type(c_ptr), value :: ctx, f
      character(kind=c_char), intent(in) :: space(*)
      real(c_double), value :: ln
      integer(c_int32_t), intent(in) :: offsets(3)
    end function
    integer(c_int) function mlx_field_get_meta(ctx, f, space4, ln, offsets) bind(C, name='mlx_field_get_meta')
      import :: c_int, c_ptr, c_char, c_double, c_int32_t
      type(c_ptr), value :: ctx, f
      character(kind=c_char), intent(out) :: space4(4)
      real(c_double), intent(out) :: ln
      integer(c_int32_t), intent(out) :: offsets(3)
    end function
    integer(c_int) function mlx_chop_offset(ctx, f, iof1, iof2, iof3) bind(C, name='mlx_chop_offset')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx, f
      integer(c_int32_t), value :: iof1, iof2, iof3
    end function
    integer(c_int) function mlx_field_axpby(ctx, dst, a, src, b) bind(C, name='mlx_field_axpby')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, dst, src
      real(c_double), value :: a, b
    end function
    integer(c_int) function mlx_trans(ctx, f, space) bind(C, name='mlx_trans')
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: ctx, f
      character(kind=c_char), intent(in) :: space(*)
    end function
    integer(c_int) function mlx_svv_filter(ctx, f, gain) bind(C, name='mlx_svv_filter')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, f
      real(c_double), intent(inout) :: gain
    end function
    integer(c_int) function mlx_calcat0(ctx, f, out) bind(C, name='mlx_calcat0')
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: ctx, f
      complex(c_double_complex), intent(out) :: out(*)
    end function
    integer(c_int) function mlx_calcat1(ctx, f, out) bind(C, name='mlx_calcat1')
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: ctx, f
      complex(c_double_complex), intent(out) :: out(*)
    end function
    integer(c_int) function mlx_idel2_preln(ctx, f, preln) bind(C, name='mlx_idel2_preln')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, f
      real(c_double), value :: preln
    end function
    integer(c_int) function mlx_ihelm(ctx, f, alpha) bind(C, name='mlx_ihelm')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, f
      real(c_double), value :: alpha
    end function
    integer(c_int) function mlx_helmp(ctx, f, power, alpha, beta) bind(C, name='mlx_helmp')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, f
      integer(c_int), value :: power
      real(c_double), value :: alpha, beta
    end function
    integer(c_int) function mlx_ihelmp(ctx, f, power, alpha, beta) bind(C, name='mlx_ihelmp')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, f
      integer(c_int), value :: power
      real(c_double), value :: alpha, beta
    end function
    integer(c_int) function mlx_fefe(ctx, s, nl, dt) bind(C, name='mlx_fefe')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, s, nl
      real(c_double), value :: dt
    end function
    integer(c_int) function mlx_abab(ctx, s, s_p, nl, nl_p, dt, is_2nd) bind(C, name='mlx_abab')
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx, s, s_p, nl, nl_p
      real(c_double), value :: dt
      integer(c_int32_t), value :: is_2nd
    end function
    integer(c_int) function mlx_febe(ctx, s, nl, dt) bind(C, name='mlx_febe')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, s, nl
      real(c_double), value :: dt
    end function
    integer(c_int) function mlx_abcn(ctx, s, s_p, nl, nl_p, dt) bind(C, name='mlx_abcn')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, s, s_p, nl, nl_p
      real(c_double), value :: dt
    end function
    integer(c_int) function mlx_check_stability(ctx, psi, chi) bind(C, name='mlx_check_stability')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, psi, chi
    end function
    integer(c_int) function mlx_energy(ctx, vr, vp, vz, out) bind(C, name='mlx_energy')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, vr, vp, vz
      real(c_double), intent(out) :: out
    end function
    integer(c_int) function mlx_vorticity_magnitude(ctx, psi, chi, out) bind(C, name='mlx_vorticity_magnitude')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, psi, chi, out
    end function
    integer(c_int) function mlx_field_save(ctx, f, path, is_binary, is_global) bind(C, name='mlx_field_save')
      import :: c_int, c_ptr, c_char, c_int32_t
      type(c_ptr), value :: ctx, f
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int32_t), value :: is_binary, is_global
    end function
    integer(c_int) function mlx_field_load(ctx, f, path, is_binary, is_global) bind(C, name='mlx_field_load')
      import :: c_int, c_ptr, c_char, c_int32_t
      type(c_ptr), value :: ctx, f
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int32_t), value :: is_binary, is_global
    end function
  end interface

  !> the one-field operators share the signature int f(ctx, field)
  abstract interface
    integer(c_int) function op1_t(ctx, f) bind(C)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, f
    end function
  end interface

  interface
    integer(c_int) function c_mlx_chop(ctx, f) bind(C, name='mlx_chop')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, f
    end function
    integer(c_int) function c_mlx_dealias(ctx, f) bind(C, name='mlx_dealias')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, f
    end function
    integer(c_int) function c_mlx_zeroat1(ctx, f) bind(C, name='mlx_zeroat1')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, f
    end function
    integer(c_int) function c_mlx_fftreat(ctx, f) bind(C, name='mlx_fftreat')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, f
    end function
    integer(c_int) function c_mlx_delsqp(ctx, f) bind(C, name='mlx_delsqp')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, f
    end function
    integer(c_int) function c_mlx_idelsqp(ctx, f) bind(C, name='mlx_idelsqp')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, f
    end function
    integer(c_int) function c_mlx_xxdx(ctx, f) bind(C, name='mlx_xxdx')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, f
    end function
    integer(c_int) function c_mlx_del2h(ctx, f) bind(C, name='mlx_del2h')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, f
    end function
    integer(c_int) function c_mlx_del2(ctx, f) bind(C, name='mlx_del2')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, f
    end function
    integer(c_int) function c_mlx_idel2(ctx, f) bind(C, name='mlx_idel2')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, f
    end function
    integer(c_int) function c_mlx_vecprod(ctx, a, b, c, d, e, f) bind(C, name='mlx_vecprod')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, a, b, c, d, e, f
    end function
    integer(c_int) function c_mlx_vec2tp(ctx, a, b, c, d, e) bind(C, name='mlx_vec2tp')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, a, b, c, d, e
    end function
    integer(c_int) function c_mlx_tp2vec(ctx, a, b, c, d, e) bind(C, name='mlx_tp2vec')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, a, b, c, d, e
    end function
    integer(c_int) function c_mlx_tp2curlvec(ctx, a, b, c, d, e) bind(C, name='mlx_tp2curlvec')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, a, b, c, d, e
    end function
    integer(c_int) function c_mlx_advection_rhs(ctx, a, b, c, d, e) bind(C, name='mlx_advection_rhs')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, a, b, c, d, e
    end function
  end interface

contains

  ! --------------------------------------------------------------------------
  ! errors
  ! --------------------------------------------------------------------------
  subroutine check(ctx, ierr, where)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: ierr
    character(len=*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    character(len=512) :: text
    integer :: i
    if (ierr .eq. 0) return
    text = ''
    call c_f_pointer(mlx_last_error(ctx), msg, [512])
    do i = 1, 512
      if (msg(i) .eq. c_null_char) exit
      text(i:i) = msg(i)
    enddo
    write(error_unit, '(a,a,a,i0,a,a)') 'ERROR: ', where, ' (code ', ierr, '): ', trim(text)
    error stop 1
  end subroutine

  function cstr(s) result(c)
    character(len=*), intent(in) :: s
    character(kind=c_char, len=len_trim(s) + 1) :: c
    c = trim(s) // c_null_char
  end function

  subroutine need_set(t, where)
    class(tfm_kit_3d), intent(in) :: t
    character(len=*), intent(in) :: where
    if (.not. t%is_set) then
      write(error_unit, '(a,a)') where, ': transformation kit is not initialized'
      error stop 1
    endif
  end subroutine

  subroutine need_field(s, where)
    class(scalar), intent(in) :: s
    character(len=*), intent(in) :: where
    if (.not. c_associated(s%h)) then
      write(error_unit, '(a,a)') where, ': scalar is not initialized'
      error stop 1
    endif
  end subroutine

  ! --------------------------------------------------------------------------
  ! input file and stepping (mlegs_misc_io.f90:307-408)
  ! --------------------------------------------------------------------------
  subroutine read_input(fn)
    character(len=*), intent(in) :: fn
    integer(i4) :: fu, ie
    logical :: xpi_to_zlen
    character(len=1) :: dum
    open(newunit=fu, file=trim(adjustl(fn)), status='old', action='read', form='formatted', iostat=ie)
    if (ie .ne. 0) error stop 'read_input: could not open input file'
    read(fu, *) dum
    read(fu, *) dum; read(fu, *) nr, np, nz
    read(fu, *) dum; read(fu, *) nrchop, npchop, nzchop
    read(fu, *) dum; read(fu, *) ell, zlen, xpi_to_zlen
    if (xpi_to_zlen) zlen = zlen * pi
    read(fu, *) dum
    read(fu, *) dum
    read(fu, *) dum; read(fu, *) dt, ti, totaltime, ni, totaln
    read(fu, *) dum
    read(fu, *) dum
    read(fu, *) dum; read(fu, *) visc, hyperpow, hypervisc
    read(fu, *) dum
    read(fu, *) dum
    read(fu, *) dum; read(fu, '(a256)') flddir
    call out_dir(flddir, 'field')
    read(fu, *) dum; read(fu, *) isfldsav, fldsavintvl
    read(fu, *) dum
    read(fu, *) dum
    read(fu, *) dum; read(fu, '(a256)') logdir
    call out_dir(logdir, 'log')
    read(fu, *) dum; read(fu, *) islogsav, logsavintvl
    read(fu, *, iostat=ie) dum
    if (ie .eq. 0) then  ! optional stability controls
      read(fu, *, iostat=ie) dum
      if (ie .eq. 0) then
        read(fu, *, iostat=ie) is_svv, svv_cutoff, svv_target, svv_strength, svv_relax
        if (ie .ne. 0) error stop 'read_input: invalid SVV controls'
      endif
    endif
    close(fu)
    if ((nr .le. 0) .or. (mod(nr, 2) .ne. 0) .or. (np .le. 0) .or. (nz .le. 0)) &
      error stop 'read_input: nr must be positive and even; np and nz must be positive'
    if ((nrchop .le. 0) .or. (nrchop .gt. nr) .or. (npchop .le. 0) .or. (npchop .gt. np/2 + 1) .or. &
        (nzchop .le. 0) .or. (nzchop .gt. nz/2 + 1)) error stop 'read_input: invalid chopping sizes'
    if ((ell .le. 0.0_p8) .or. (zlen .le. 0.0_p8)) error stop 'read_input: ell and zlen must be positive'
    if ((dt .le. 0.0_p8) .or. (totaltime .lt. 0.0_p8) .or. (totaln .lt. 0)) &
      error stop 'read_input: invalid time stepping values'
    if ((visc .lt. 0.0_p8) .or. (hypervisc .lt. 0.0_p8)) error stop 'read_input: viscosities must be non-negative'
    if ((hyperpow .ne. 0) .and. (hyperpow .ne. 4) .and. (hyperpow .ne. 6) .and. (hyperpow .ne. 8)) &
      error stop 'read_input: hyperpow must be zero, 4, 6, or 8'
    if ((hyperpow .ne. 0) .and. (hypervisc .le. 0.0_p8)) error stop 'read_input: hypervisc must be positive when enabled'
    if ((fldsavintvl .le. 0) .or. (logsavintvl .le. 0)) error stop 'read_input: output intervals must be positive'
    if ((svv_cutoff .le. 0.0_p8) .or. (svv_cutoff .ge. 1.0_p8) .or. (svv_target .le. 0.0_p8) .or. &
        (svv_strength .lt. 0.0_p8) .or. (svv_relax .lt. 0.0_p8) .or. (svv_relax .gt. 1.0_p8)) &
      error stop 'read_input: invalid SVV controls'
  end subroutine

  subroutine out_dir(d, what)
    character(len=*), intent(inout) :: d
    character(len=*), intent(in) :: what
    integer :: cmdstat, exitstat
    d = trim(adjustl(d))
    if (d(len_trim(d):len_trim(d)) .ne. '/') d = trim(d) // '/'
    if (verify(trim(d), 'abcdefghijklmnopqrstuvwxyzABCDEFGHIJKLMNOPQRSTUVWXYZ0123456789_./-') .ne. 0) then
      write(error_unit, '(a,a,a)') 'read_input: invalid ', what, ' output directory'
      error stop 1
    endif
    call execute_command_line('mkdir -p -- "' // trim(d) // '"', wait=.true., exitstat=exitstat, cmdstat=cmdstat)
    if ((cmdstat .ne. 0) .or. (exitstat .ne. 0)) then
      write(error_unit, '(a,a,a)') 'read_input: could not create ', what, ' output directory'
      error stop 1
    endif
  end subroutine

  subroutine timestep_set()
    real(p8) :: tott
    integer(i4) :: totn
    if (dt .le. 0.0_p8) error stop 'timestep_setup: dt must be positive'
    tott = totaltime
    totn = totaln
    totaltime = max(tott, dt * totn)
    if (totaltime .eq. tott) then
      totaln = floor(tott / dt)
    else
      totaln = totn
    endif
    curr_n = ni
    curr_t = ti
  end subroutine

  ! --------------------------------------------------------------------------
  ! transformation kit
  ! --------------------------------------------------------------------------
  subroutine kit_init(this)
    class(tfm_kit_3d), intent(inout) :: this
    type(mlx_params_c) :: p
    integer(c_int32_t) :: d(9)
    integer(i4) :: i
    if (this%is_set) call this%dealloc()
    p%nr = nr; p%np = np; p%nz = nz
    p%nrchop = nrchop; p%npchop = npchop; p%nzchop = nzchop
    p%ell = ell; p%zlen = zlen; p%dt = dt; p%ti = ti; p%totaltime = totaltime
    p%ni = ni; p%totaln = totaln; p%visc = visc; p%hyperpow = hyperpow; p%hypervisc = hypervisc
    p%is_svv = merge(1, 0, is_svv)
    p%svv_cutoff = svv_cutoff; p%svv_target = svv_target; p%svv_strength = svv_strength; p%svv_relax = svv_relax
    call check(c_null_ptr, mlx_ctx_create(p, int(mlx_device, c_int), c_null_ptr, this%ctx), 'tfm%init')
    call check(this%ctx, mlx_kit_init(this%ctx), 'tfm%init')
    call check(this%ctx, mlx_kit_dims(this%ctx, d), 'tfm%init')
    this%nrdim = d(1); this%npdim = d(2); this%nzdim = d(3); this%nrh = d(4); this%ne = d(5)
    this%chopzl = d(6); this%chopzu = d(7); this%nph = d(8); this%chopp = d(9)
    allocate(this%x(nr), this%w(nr), this%ln(nr), this%r(nr), this%ak(this%nzdim))
    call check(this%ctx, mlx_kit_get(this%ctx, cstr('x'), this%x, int(nr, c_int64_t)), 'tfm%init')
    call check(this%ctx, mlx_kit_get(this%ctx, cstr('w'), this%w, int(nr, c_int64_t)), 'tfm%init')
    call check(this%ctx, mlx_kit_get(this%ctx, cstr('ln'), this%ln, int(nr, c_int64_t)), 'tfm%init')
    call check(this%ctx, mlx_kit_get(this%ctx, cstr('r'), this%r, int(nr, c_int64_t)), 'tfm%init')
    call check(this%ctx, mlx_kit_get(this%ctx, cstr('ak'), this%ak, int(this%nzdim, c_int64_t)), 'tfm%init')
    allocate(this%p(np + 1), this%z(nz + 1))
    this%p = [(2.0_p8 * pi / np * i, i = 0, np)]      ! mlegs_spectfm_init.f90:119-120
    this%z = [(zlen / nz * i, i = 0, nz)]              ! :129-130
    call check(this%ctx, mlx_set_precision(this%ctx, int(mlx_precision, c_int)), 'tfm%init')
    this%is_set = .true.
  end subroutine

  subroutine kit_dealloc(this)
    class(tfm_kit_3d), intent(inout) :: this
    integer(c_int) :: ierr
    if (c_associated(this%ctx)) ierr = mlx_ctx_destroy(this%ctx)
    this%ctx = c_null_ptr
    if (allocated(this%x)) deallocate(this%x, this%w, this%ln, this%r, this%ak, this%p, this%z)
    this%is_set = .false.
  end subroutine

  subroutine kit_set_precision(this, prec)
    class(tfm_kit_3d), intent(inout) :: this
    integer(i4), intent(in) :: prec
    call need_set(this, 'set_precision')
    call check(this%ctx, mlx_set_precision(this%ctx, int(prec, c_int)), 'set_precision')
  end subroutine

  !> precision of one contraction site (MLX_STAGE_*); prec = -1 follows set_precision
  subroutine kit_set_stage_precision(this, stage, prec)
    class(tfm_kit_3d), intent(inout) :: this
    integer(i4), intent(in) :: stage, prec
    call need_set(this, 'set_stage_precision')
    call check(this%ctx, mlx_set_stage_precision(this%ctx, int(stage, c_int), int(prec, c_int)), &
               'set_stage_precision')
  end subroutine

  subroutine kit_sync(this)
    class(tfm_kit_3d), intent(inout) :: this
    call check(this%ctx, mlx_sync(this%ctx), 'sync')
  end subroutine

  ! --------------------------------------------------------------------------
  ! scalar lifetime, metadata and host access (mlegs_scalar_init.f90)
  ! --------------------------------------------------------------------------
  subroutine scalar_init(this, glb_sz, axis_comm)
    class(scalar), intent(inout) :: this
    integer(i4), intent(in) :: glb_sz(3)
    integer(i4), intent(in) :: axis_comm(:)
    call need_set(tfm, 'scalar_init')
    if (any(glb_sz .ne. [tfm%nrdim, tfm%npdim, tfm%nzdim])) &
      error stop 'scalar_init: glb_sz must equal (nrdim, npdim, nzdim) of the kit'
    if (size(axis_comm) .ne. 3) error stop 'scalar_init: axis_comm must have three entries'
    if (c_associated(this%h)) call this%dealloc()
    this%ctx = tfm%ctx
    this%glb_sz = glb_sz
    call check(this%ctx, mlx_field_create(this%ctx, this%h), 'scalar_init')
  end subroutine

  subroutine scalar_dealloc(this)
    class(scalar), intent(inout) :: this
    integer(c_int) :: ierr
    if (c_associated(this%h)) ierr = mlx_field_destroy(this%ctx, this%h)
    this%h = c_null_ptr
  end subroutine

  subroutine scalar_chop_offset(this, iof1, iof2, iof3)
    class(scalar), intent(inout) :: this
    integer(i4), intent(in) :: iof1
    integer(i4), optional :: iof2, iof3
    integer(c_int32_t) :: o2, o3
    o2 = 0; o3 = 0
    if (present(iof2)) o2 = iof2
    if (present(iof3)) o3 = iof3
    call need_field(this, 'chop_offset')
    call check(this%ctx, mlx_chop_offset(this%ctx, this%h, int(iof1, c_int32_t), o2, o3), 'chop_offset')
  end subroutine

  subroutine meta(this, space, ln, offsets)
    class(scalar), intent(in) :: this
    character(len=3), intent(out) :: space
    real(p8), intent(out) :: ln
    integer(c_int32_t), intent(out) :: offsets(3)
    character(kind=c_char) :: s4(4)
    call check(this%ctx, mlx_field_get_meta(this%ctx, this%h, s4, ln, offsets), 'scalar meta')
    space = s4(1) // s4(2) // s4(3)
  end subroutine

  function scalar_get_space(this) result(space)
    class(scalar), intent(in) :: this
    character(len=3) :: space
    real(p8) :: ln
    integer(c_int32_t) :: off(3)
    call need_field(this, 'get_space')
    call meta(this, space, ln, off)
  end function

  subroutine scalar_set_space(this, space)
    class(scalar), intent(inout) :: this
    character(len=3), intent(in) :: space
    character(len=3) :: sp
    real(p8) :: ln
    integer(c_int32_t) :: off(3)
    call need_field(this, 'set_space')
    call meta(this, sp, ln, off)
    call check(this%ctx, mlx_field_set_meta(this%ctx, this%h, cstr(space), ln, off), 'set_space')
  end subroutine

  real(p8) function scalar_get_ln(this) result(ln)
    class(scalar), intent(in) :: this
    character(len=3) :: sp
    integer(c_int32_t) :: off(3)
    call need_field(this, 'get_ln')
    call meta(this, sp, ln, off)
  end function

  subroutine scalar_set_ln(this, ln)
    class(scalar), intent(inout) :: this
    real(p8), intent(in) :: ln
    character(len=3) :: sp
    real(p8) :: old
    integer(c_int32_t) :: off(3)
    call need_field(this, 'set_ln')
    call meta(this, sp, old, off)
    call check(this%ctx, mlx_field_set_meta(this%ctx, this%h, cstr(sp), ln, off), 'set_ln')
  end subroutine

  !> the whole field as a global array e(nrdim, npdim, nzdim)
  function scalar_assemble(this, axis_input) result(array_glb)
    class(scalar), intent(in) :: this
    integer(i4), optional :: axis_input
    complex(p8), dimension(:,:,:), allocatable :: array_glb
    character(kind=c_char) :: s4(4)
    real(c_double) :: ln
    integer(c_int32_t) :: off(3)
    call need_field(this, 'assemble')
    allocate(array_glb(this%glb_sz(1), this%glb_sz(2), this%glb_sz(3)))
    call check(this%ctx, mlx_field_download(this%ctx, this%h, array_glb, int(this%glb_sz(1), c_int64_t), &
               int(this%glb_sz(2), c_int64_t), int(this%glb_sz(3), c_int64_t), s4, ln, off), 'assemble')
  end function

  !> store a global array; space, ln and offsets are kept
  subroutine scalar_disassemble(this, array_glb, axis_input)
    class(scalar), intent(inout) :: this
    complex(p8), dimension(:,:,:), allocatable :: array_glb
    integer(i4), optional :: axis_input
    character(len=3) :: sp
    real(p8) :: ln
    integer(c_int32_t) :: off(3)
    call need_field(this, 'disassemble')
    if (any(shape(array_glb) .ne. this%glb_sz)) error stop 'disassemble: array shape must equal glb_sz'
    call meta(this, sp, ln, off)
    call check(this%ctx, mlx_field_upload(this%ctx, this%h, array_glb, int(this%glb_sz(1), c_int64_t), &
               int(this%glb_sz(2), c_int64_t), int(this%glb_sz(3), c_int64_t), cstr(sp), ln, off), 'disassemble')
  end subroutine

  !> this = a*this + b*that on data and ln (the drivers' s%e = a*s%e + b*t%e)
  subroutine scalar_axpby(this, a, that, b)
    class(scalar), intent(inout) :: this
    real(p8), intent(in) :: a, b
    class(scalar), intent(in) :: that
    call check(this%ctx, mlx_field_axpby(this%ctx, this%h, a, that%h, b), 'axpby')
  end subroutine

  !> s%e = 0 of the reference: the data zeroed, space, log term and offsets kept
  subroutine scalar_zero_e(this)
    class(scalar), intent(inout) :: this
    complex(p8), dimension(:,:,:), allocatable :: z
    allocate(z(this%glb_sz(1), this%glb_sz(2), this%glb_sz(3)))
    z = (0.0_p8, 0.0_p8)
    call this%disassemble(z)
    deallocate(z)
  end subroutine

  !> deep copy this = that (mlegs_scalar_init.f90:106-140)
  subroutine scalar_copy(this, that)
    class(scalar), intent(inout) :: this
    class(scalar), intent(in) :: that
    call need_field(that, 'scalar_copy')
    if (.not. c_associated(this%h)) then
      this%ctx = that%ctx
      this%glb_sz = that%glb_sz
      call check(this%ctx, mlx_field_create(this%ctx, this%h), 'scalar_copy')
    endif
    call check(this%ctx, mlx_field_copy(this%ctx, this%h, that%h), 'scalar_copy')
  end subroutine

  ! --------------------------------------------------------------------------
  ! operators (mlegs_scalar.f90:116-428)
  ! --------------------------------------------------------------------------
  subroutine scalar_trans(s, space, tfm)
    class(scalar), intent(inout) :: s
    character(len=3), intent(in) :: space
    class(tfm_kit_3d), intent(in) :: tfm
    call need_set(tfm, 'trans')
    call check(tfm%ctx, mlx_trans(tfm%ctx, s%h, cstr(space)), 'trans')
  end subroutine

  subroutine one(fn, s, tfm, where)
    procedure(op1_t) :: fn
    class(scalar), intent(inout) :: s
    class(tfm_kit_3d), intent(in) :: tfm
    character(len=*), intent(in) :: where
    call need_set(tfm, where)
    call need_field(s, where)
    call check(tfm%ctx, fn(tfm%ctx, s%h), where)
  end subroutine

  subroutine scalar_chop(s, tfm)
    class(scalar), intent(inout) :: s
    class(tfm_kit_3d), intent(in) :: tfm
    call one(c_mlx_chop, s, tfm, 'chop')
  end subroutine

  subroutine scalar_dealias(s, tfm)
    class(scalar), intent(inout) :: s
    class(tfm_kit_3d), intent(in) :: tfm
    call one(c_mlx_dealias, s, tfm, 'dealias')
  end subroutine

  subroutine scalar_svv_filter(s, tfm, gain)
    class(scalar), intent(inout) :: s
    class(tfm_kit_3d), intent(in) :: tfm
    real(p8), intent(inout) :: gain
    call need_set(tfm, 'svv_filter')
    call check(tfm%ctx, mlx_svv_filter(tfm%ctx, s%h, gain), 'svv_filter')
  end subroutine

  function scalar_calcat0(s, tfm) result(calc)
    class(scalar) :: s
    class(tfm_kit_3d) :: tfm
    complex(p8), dimension(:), allocatable :: calc
    call need_set(tfm, 'calcat0')
    allocate(calc(tfm%nzdim))
    call check(tfm%ctx, mlx_calcat0(tfm%ctx, s%h, calc), 'calcat0')
  end function

  function scalar_calcat1(s, tfm) result(calc)
    class(scalar) :: s
    class(tfm_kit_3d) :: tfm
    complex(p8), dimension(:), allocatable :: calc
    call need_set(tfm, 'calcat1')
    allocate(calc(tfm%nzdim))
    call check(tfm%ctx, mlx_calcat1(tfm%ctx, s%h, calc), 'calcat1')
  end function

  subroutine scalar_zeroat1(s, tfm)
    class(scalar), intent(inout) :: s
    class(tfm_kit_3d), intent(in) :: tfm
    call one(c_mlx_zeroat1, s, tfm, 'zeroat1')
  end subroutine

  !> fftreat(s, tfm) (mlegs_scalar_ops.f90:1002-1063): far-field smoothing
  subroutine scalar_fftreat(s, tfm)
    class(scalar), intent(inout) :: s
    class(tfm_kit_3d), intent(in) :: tfm
    call one(c_mlx_fftreat, s, tfm, 'fftreat')
  end subroutine

  subroutine scalar_delsqp(s, tfm)
    class(scalar), intent(inout) :: s
    class(tfm_kit_3d) :: tfm
    call one(c_mlx_delsqp, s, tfm, 'delsqp')
  end subroutine

  subroutine scalar_idelsqp(s, tfm)
    class(scalar), intent(inout) :: s
    class(tfm_kit_3d) :: tfm
    call one(c_mlx_idelsqp, s, tfm, 'idelsqp')
  end subroutine

  subroutine scalar_xxdx(s, tfm)
    class(scalar), intent(inout) :: s
    class(tfm_kit_3d) :: tfm
    call one(c_mlx_xxdx, s, tfm, 'xxdx')
  end subroutine

  subroutine scalar_del2h(s, tfm)
    class(scalar), intent(inout) :: s
    class(tfm_kit_3d) :: tfm
    call one(c_mlx_del2h, s, tfm, 'del2h')
  end subroutine

  subroutine scalar_del2(s, tfm)
    class(scalar), intent(inout) :: s
    class(tfm_kit_3d) :: tfm
    call one(c_mlx_del2, s, tfm, 'del2')
  end subroutine

  subroutine scalar_idel2_proln(s, tfm)
    class(scalar), intent(inout) :: s
    class(tfm_kit_3d) :: tfm
    call one(c_mlx_idel2, s, tfm, 'idel2')
  end subroutine

  subroutine scalar_idel2_preln(s, tfm, preln)
    class(scalar), intent(inout) :: s
    class(tfm_kit_3d) :: tfm
    real(p8) :: preln
    call need_set(tfm, 'idel2')
    call check(tfm%ctx, mlx_idel2_preln(tfm%ctx, s%h, preln), 'idel2')
  end subroutine

  subroutine scalar_ihelm(s, alpha, tfm)
    class(scalar), intent(inout) :: s
    real(p8) :: alpha
    class(tfm_kit_3d) :: tfm
    call need_set(tfm, 'ihelm')
    call check(tfm%ctx, mlx_ihelm(tfm%ctx, s%h, alpha), 'ihelm')
  end subroutine

  ! helmp(s, power, alpha, beta, tfm) (mlegs_scalar.f90:288-297)
  subroutine scalar_helmp(s, power, alpha, beta, tfm)
    class(scalar), intent(inout) :: s
    integer(i4) :: power
    real(p8) :: alpha, beta
    class(tfm_kit_3d) :: tfm
    call need_set(tfm, 'helmp')
    call check(tfm%ctx, mlx_helmp(tfm%ctx, s%h, int(power, c_int), alpha, beta), 'helmp')
  end subroutine

  ! ihelmp(s, power, alpha, beta, tfm) (mlegs_scalar.f90:300-309)
  subroutine scalar_ihelmp(s, power, alpha, beta, tfm)
    class(scalar), intent(inout) :: s
    integer(i4) :: power
    real(p8) :: alpha, beta
    class(tfm_kit_3d) :: tfm
    call need_set(tfm, 'ihelmp')
    call check(tfm%ctx, mlx_ihelmp(tfm%ctx, s%h, int(power, c_int), alpha, beta), 'ihelmp')
  end subroutine

  ! fefe(s, s_rhs_nonlin, dt, tfm) (mlegs_scalar.f90:316-325)
  subroutine scalar_fefe(s, s_rhs_nonlin, dt, tfm)
    class(scalar), intent(inout) :: s
    class(scalar), intent(in) :: s_rhs_nonlin
    real(p8), intent(in) :: dt
    class(tfm_kit_3d) :: tfm
    call need_set(tfm, 'fefe')
    call check(tfm%ctx, mlx_fefe(tfm%ctx, s%h, s_rhs_nonlin%h, dt), 'fefe')
  end subroutine

  ! abab(s, s_p, s_rhs_nonlin, s_rhs_nonlin_p, dt, tfm, is_2nd_svis_p) (mlegs_scalar.f90:332-342)
  subroutine scalar_abab(s, s_p, s_rhs_nonlin, s_rhs_nonlin_p, dt, tfm, is_2nd_svis_p)
    class(scalar), intent(inout) :: s, s_rhs_nonlin
    class(scalar), intent(inout) :: s_p, s_rhs_nonlin_p
    real(p8), intent(in) :: dt
    class(tfm_kit_3d) :: tfm
    logical, optional :: is_2nd_svis_p
    integer(c_int32_t) :: flag
    flag = 0
    if (present(is_2nd_svis_p)) then
      if (is_2nd_svis_p) flag = 1
    endif
    call need_set(tfm, 'abab')
    call check(tfm%ctx, mlx_abab(tfm%ctx, s%h, s_p%h, s_rhs_nonlin%h, s_rhs_nonlin_p%h, dt, flag), 'abab')
  end subroutine

  subroutine scalar_febe(s, s_rhs_nonlin, dt, tfm)
    class(scalar), intent(inout) :: s
    class(scalar), intent(in) :: s_rhs_nonlin
    real(p8), intent(in) :: dt
    class(tfm_kit_3d) :: tfm
    call need_set(tfm, 'febe')
    call check(tfm%ctx, mlx_febe(tfm%ctx, s%h, s_rhs_nonlin%h, dt), 'febe')
  end subroutine

  subroutine scalar_abcn(s, s_p, s_rhs_nonlin, s_rhs_nonlin_p, dt, tfm)
    class(scalar), intent(inout) :: s, s_rhs_nonlin
    class(scalar), intent(inout) :: s_p, s_rhs_nonlin_p
    real(p8), intent(in) :: dt
    class(tfm_kit_3d) :: tfm
    call need_set(tfm, 'abcn')
    call check(tfm%ctx, mlx_abcn(tfm%ctx, s%h, s_p%h, s_rhs_nonlin%h, s_rhs_nonlin_p%h, dt), 'abcn')
  end subroutine

  subroutine vector_product(vr, vp, vz, ur, up, uz, tfm)
    class(scalar), intent(inout) :: vr, vp, vz
    class(scalar), intent(in) :: ur, up, uz
    class(tfm_kit_3d) :: tfm
    call need_set(tfm, 'vecprod')
    call check(tfm%ctx, c_mlx_vecprod(tfm%ctx, vr%h, vp%h, vz%h, ur%h, up%h, uz%h), 'vecprod')
  end subroutine

  subroutine vector_projection(vr, vp, vz, psi, chi, tfm)
    class(scalar), intent(in) :: vr, vp, vz
    class(scalar), intent(inout) :: psi, chi
    class(tfm_kit_3d) :: tfm
    call need_set(tfm, 'vec2tp')
    call check(tfm%ctx, c_mlx_vec2tp(tfm%ctx, vr%h, vp%h, vz%h, psi%h, chi%h), 'vec2tp')
  end subroutine

  subroutine vector_reconstruction(psi, chi, vr, vp, vz, tfm)
    class(scalar), intent(in) :: psi, chi
    class(scalar), intent(inout) :: vr, vp, vz
    class(tfm_kit_3d) :: tfm
    call need_set(tfm, 'tp2vec')
    call check(tfm%ctx, c_mlx_tp2vec(tfm%ctx, psi%h, chi%h, vr%h, vp%h, vz%h), 'tp2vec')
  end subroutine

  subroutine curl_vector_reconstruction(psi, chi, wr, wp, wz, tfm)
    class(scalar), intent(in) :: psi, chi
    class(scalar), intent(inout) :: wr, wp, wz
    class(tfm_kit_3d) :: tfm
    call need_set(tfm, 'tp2curlvec')
    call check(tfm%ctx, c_mlx_tp2curlvec(tfm%ctx, psi%h, chi%h, wr%h, wp%h, wz%h), 'tp2curlvec')
  end subroutine

  !> N = P_TP[(U + V) x curl V], de-aliased before and after; psi and chi are
  !> de-aliased in place (vortical_flow_3d.f90:353-395, on the device)
  subroutine advection_rhs(psi, chi, nlpsi, nlchi, uz)
    type(scalar), intent(inout) :: psi, chi
    type(scalar), intent(in) :: uz
    type(scalar), intent(inout) :: nlpsi, nlchi
    call need_set(tfm, 'advection_rhs')
    call check(tfm%ctx, c_mlx_advection_rhs(tfm%ctx, psi%h, chi%h, nlpsi%h, nlchi%h, uz%h), 'advection_rhs')
  end subroutine

  !> stops the program on a non-finite state (vortical_flow_3d.f90:397-409)
  subroutine check_stability(psi, chi)
    type(scalar), intent(in) :: psi, chi
    integer(c_int) :: ierr
    ierr = mlx_check_stability(tfm%ctx, psi%h, chi%h)
    if (ierr .ne. 0) then
      write(error_unit, '(a,i0)') 'ERROR: non-finite vortex state at time step ', curr_n
      error stop 205
    endif
  end subroutine

  !> 1/2 of the quadrature of |u|^2 over the domain (FP64, on the device)
  real(p8) function energy(vr, vp, vz) result(e)
    type(scalar), intent(in) :: vr, vp, vz
    call check(tfm%ctx, mlx_energy(tfm%ctx, vr%h, vp%h, vz%h, e), 'energy')
  end function

  !> |curl u| of (psi, chi) at each point, real and imaginary parts apart:
  !> the field of save_vort_mag (vortical_flow_3d.f90:411-446), on the device
  subroutine vorticity_magnitude(psi, chi, vormag)
    type(scalar), intent(in) :: psi, chi
    type(scalar), intent(inout) :: vormag
    call need_set(tfm, 'vorticity_magnitude')
    call check(tfm%ctx, mlx_vorticity_magnitude(tfm%ctx, psi%h, chi%h, vormag%h), 'vorticity_magnitude')
  end subroutine

  !> msave(s, fn, is_binary, is_global) (mlegs_scalar_io.f90:6-101): ASCII
  !> unless is_binary; one global file unless is_global is .false. (then
  !> fn_0, this process's block).  Defaults as the reference documents them.
  subroutine msave_scalar(s, fn, is_binary, is_global)
    type(scalar), intent(in) :: s
    character(len=*), intent(in) :: fn
    logical, intent(in), optional :: is_binary, is_global
    integer(c_int32_t) :: b, g
    b = 0; g = 1
    if (present(is_binary)) then
      if (is_binary) b = 1
    endif
    if (present(is_global)) then
      if (.not. is_global) g = 0
    endif
    call check(s%ctx, mlx_field_save(s%ctx, s%h, trim(fn) // c_null_char, b, g), 'msave')
  end subroutine

  !> the collocation points r, phi (np > 1) and z (nz > 1) written to
  !> <logdir>coords_{r,p,z}.dat, the colloc_pts routine of the reference's
  !> programs (e.g. src/apps/scalar_diffusion_2d.f90, contains section)
  subroutine colloc_pts(kit)
    class(tfm_kit_3d), intent(in) :: kit
    call need_set(kit, 'colloc_pts')
    call msave(kit%r, trim(adjustl(logdir)) // 'coords_r.dat')
    if (np .gt. 1) call msave(kit%p, trim(adjustl(logdir)) // 'coords_p.dat')
    if (nz .gt. 1) call msave(kit%z, trim(adjustl(logdir)) // 'coords_z.dat')
  end subroutine

  !> msave of a real vector (mlegs_misc_io.f90:37-48, 484-520): the sizes
  !> n 1 1 as three I10 fields, then one value per line in 1PE24.15E3
  !> (binary: three int32 sizes, then the values)
  subroutine msave_real1(a, fn, is_binary)
    real(p8), dimension(:), intent(in) :: a
    character(len=*), intent(in) :: fn
    logical, intent(in), optional :: is_binary
    integer :: fo, i
    logical :: b
    b = .false.
    if (present(is_binary)) b = is_binary
    if (b) then
      open(newunit=fo, file=fn, form='unformatted', access='stream', status='replace')
      write(fo) int(size(a), i4), 1_i4, 1_i4
      write(fo) a
    else
      open(newunit=fo, file=fn, status='unknown')
      write(unit=fo, fmt='(3(1X,I10))', advance='no') size(a), 1, 1
      write(fo, *) ''
      do i = 1, size(a)
        write(unit=fo, fmt='(1PE24.15E3)', advance='no') a(i)
        if (i .lt. size(a)) write(fo, *) ''
      enddo
    endif
    close(fo)
    write(*, '(I6, a, A)') size(a), ' x    1 1D real vector data written to ', trim(fn)
  end subroutine

  !> mload(fn, s, is_binary, is_global) (mlegs_scalar_io.f90:103-248)
  subroutine mload_scalar(fn, s, is_binary, is_global)
    character(len=*), intent(in) :: fn
    type(scalar), intent(inout) :: s
    logical, intent(in), optional :: is_binary, is_global
    integer(c_int32_t) :: b, g
    b = 0; g = 1
    if (present(is_binary)) then
      if (is_binary) b = 1
    endif
    if (present(is_global)) then
      if (.not. is_global) g = 0
    endif
    call check(s%ctx, mlx_field_load(s%ctx, s%h, trim(fn) // c_null_char, b, g), 'mload')
  end subroutine

end module mlegs_amd
