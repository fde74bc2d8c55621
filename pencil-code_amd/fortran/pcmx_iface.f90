! pcmx_iface.f90 -- iso_c_binding interface to libpcmx (include/pcmx.h).
!
! The thin Fortran layer between a Pencil Code `module GPU` (or any Fortran
! host) and the HIP engine: one bind(C) interface per C entry point, the
! bind(C) parameter record, and a helper that turns a negative return code
! into the library's message.  Arguments cross as plain values and
! pointers; the f-array is passed as the reference's f(mx,my,mz,mfarray).
!
module pcmx_iface
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: PCMX_OK = 0
  integer(c_int), parameter :: PCMX_ERR_ARG = -1, PCMX_ERR_STATE = -2
  integer(c_int), parameter :: PCMX_ERR_HIP = -3, PCMX_ERR_NOTFOUND = -4
  integer(c_int), parameter :: PCMX_PREC_FP64 = 0, PCMX_PREC_MIXED = 1
  integer(c_int), parameter :: PCMX_N_REDUCED = 20

  type, bind(C) :: pcmx_params
    integer(c_int) :: nx, ny, nz, mvar
    integer(c_int) :: iux, ilnrho, iss, iax          ! 0-based slots, -1 if absent
    real(c_double) :: dx, dy, dz
    real(c_double) :: nu, eta, chi
    real(c_double) :: cs0, gamma, cp, rho0, mu0
    real(c_double) :: cdt, cdtv
    integer(c_int) :: ldt, precision, kernel
    integer(c_int) :: lupw_uu, lupw_lnrho
    real(c_double) :: gravz
    integer(c_int) :: lperi_z
    integer(c_int) :: bcz_bot(8), bcz_top(8)
    integer(c_int) :: lupw_ss
    ! layer physics of samples/conv-slab (direct kernel)
    real(c_double) :: omega, theta
    integer(c_int) :: lheatc_kprof, cooling_profile
    real(c_double) :: hcond0, hcond1, hcond2, widthss, z1, z2
    real(c_double) :: luminosity, wheat, cool, wcool, cs2cool
    real(c_double) :: fbot, ftop, cs2bot, cs2top
    real(c_double) :: z0, lxyz(3)
  end type pcmx_params

  ! z boundary codes of pcmx_params (PCMX_BC_* of include/pcmx.h)
  integer(c_int), parameter :: PCMX_BC_S = 1, PCMX_BC_A = 2, PCMX_BC_A2 = 3
  integer(c_int), parameter :: PCMX_BC_C1 = 4, PCMX_BC_CT = 5

  ! pcmx_forcing of include/pcmx.h: helical forcing after each step
  type, bind(C) :: pcmx_forcing
    integer(c_int) :: active, target, scale_sqrt_dt, reserved
    real(c_double) :: kk(3), phase, fact
    real(c_double) :: coef1(3), coef2(3), fda(3)
    real(c_double) :: k1_ff
    real(c_double) :: xyz0(3)
  end type pcmx_forcing

  interface
    type(c_ptr) function pcmx_last_error() bind(C, name='pcmx_last_error')
      import :: c_ptr
    end function
    integer(c_int) function pcmx_register() bind(C, name='pcmx_register')
      import :: c_int
    end function
    integer(c_int) function pcmx_initialize(f, mfarray, par, t, nt) bind(C, name='pcmx_initialize')
      import :: c_int, c_double, pcmx_params
      real(c_double), intent(in) :: f(*)
      integer(c_int), value :: mfarray
      type(pcmx_params), intent(in) :: par
      real(c_double), value :: t
      integer(c_int), value :: nt
    end function
    integer(c_int) function pcmx_finalize() bind(C, name='pcmx_finalize')
      import :: c_int
    end function
    integer(c_int) function pcmx_load_farray(f, mfarray) bind(C, name='pcmx_load_farray')
      import :: c_int, c_double
      real(c_double), intent(in) :: f(*)
      integer(c_int), value :: mfarray
    end function
    integer(c_int) function pcmx_copy_farray(f, mfarray) bind(C, name='pcmx_copy_farray')
      import :: c_int, c_double
      real(c_double), intent(inout) :: f(*)
      integer(c_int), value :: mfarray
    end function
    integer(c_int) function pcmx_before_boundary(lrmv, isubstep, t) bind(C, name='pcmx_before_boundary')
      import :: c_int, c_double
      integer(c_int), value :: lrmv, isubstep
      real(c_double), value :: t
    end function
    integer(c_int) function pcmx_rhs(isubstep, t, dt) bind(C, name='pcmx_rhs')
      import :: c_int, c_double
      integer(c_int), value :: isubstep
      real(c_double), value :: t
      real(c_double), intent(inout) :: dt
    end function
    integer(c_int) function pcmx_update_after_substep() bind(C, name='pcmx_update_after_substep')
      import :: c_int
    end function
    integer(c_int) function pcmx_prepare_for_first_substep(t, dt1) bind(C, name='pcmx_prepare_for_first_substep')
      import :: c_int, c_double
      real(c_double), value :: t
      real(c_double), intent(out) :: dt1
    end function
    integer(c_int) function pcmx_get_reduced_vars(dst) bind(C, name='pcmx_get_reduced_vars')
      import :: c_int, c_double
      real(c_double), intent(out) :: dst(*)
    end function
    integer(c_int) function pcmx_update_by_name(varname, value) bind(C, name='pcmx_update_by_name')
      import :: c_int, c_double, c_char
      character(kind=c_char), intent(in) :: varname(*)
      real(c_double), value :: value
    end function
    integer(c_int) function pcmx_update_by_index(index, value) bind(C, name='pcmx_update_by_index')
      import :: c_int, c_double
      integer(c_int), value :: index
      real(c_double), value :: value
    end function
    integer(c_int) function pcmx_test_bcs(maxdiff) bind(C, name='pcmx_test_bcs')
      import :: c_int, c_double
      real(c_double), intent(out) :: maxdiff
    end function
    integer(c_int) function pcmx_split_update() bind(C, name='pcmx_split_update')
      import :: c_int
    end function
    integer(c_int) function pcmx_time_step(nsteps, t, dt, ldiag) bind(C, name='pcmx_time_step')
      import :: c_int, c_double
      integer(c_int), value :: nsteps, ldiag
      real(c_double), intent(inout) :: t, dt
    end function
    ! var.dat written from the device state (io_dist.f90:108-199 layout,
    ! lpersist=F); path NUL-terminated (trim(file)//c_null_char)
    integer(c_int) function pcmx_write_snapshot(path, t, xyz0) bind(C, name='pcmx_write_snapshot')
      import :: c_int, c_double, c_char
      character(kind=c_char), dimension(*), intent(in) :: path
      real(c_double), value :: t
      real(c_double), dimension(3), intent(in) :: xyz0
    end function
    ! times the split kernel's tile shapes and z-chunk lengths on the loaded
    ! state and keeps the fastest: best = ms per launch, tile width, z chunk
    integer(c_int) function pcmx_tune(best) bind(C, name='pcmx_tune')
      import :: c_int, c_double
      real(c_double), dimension(3), intent(out) :: best
    end function
    integer(c_int) function pcmx_set_forcing(frc) bind(C, name='pcmx_set_forcing')
      import :: c_int, pcmx_forcing
      type(pcmx_forcing), intent(in) :: frc
    end function
  end interface

contains

  ! Message of the last failed call, as a Fortran string.
  function pcmx_error_message() result(msg)
    character(len=1024) :: msg
    character(kind=c_char), pointer :: cs(:)
    type(c_ptr) :: p
    integer :: i
    msg = ''
    p = pcmx_last_error()
    if (.not. c_associated(p)) return
    call c_f_pointer(p, cs, [1024])
    do i = 1, 1024
      if (cs(i) == c_null_char) exit
      msg(i:i) = cs(i)
    enddo
  end function pcmx_error_message

end module pcmx_iface
