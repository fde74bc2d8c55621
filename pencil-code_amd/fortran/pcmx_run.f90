! pcmx_run.f90 -- a Fortran host that drives libpcmx through iso_c_binding
! in the call order of the reference time loop (timestep.f90:77-204 with the
! lgpu branches of equ.f90:192-327): register, initialize + load f, then per
! substep before_boundary -> rhs -> update_after_substep, split_update per
! step, copy f back, finalize.
!
! Usage:  pcmx_run case.nml
! The namelist &pcmx_case gives the box, the physics and two raw files: the
! initial f(mx,my,mz,mvar) (stream, real(8), Fortran order, ghosts ignored)
! and the file that receives f after nt steps (with periodic ghosts).  One
! line per step "it t dt urms umax brms" goes to standard output.
!
program pcmx_run
  use iso_c_binding
  use pcmx_iface
  implicit none
  integer :: nx = 32, ny = 32, nz = 32, nt = 1
  logical :: lmagnetic = .false., lentropy = .false., ldt = .false.
  real(8) :: Lx = 6.283185307179586d0, Ly = 6.283185307179586d0, Lz = 6.283185307179586d0
  real(8) :: nu = 5d-3, eta = 0d0, chi = 0d0, cs0 = 1d0, gamma = 1d0, cp = 1d0, rho0 = 1d0, mu0 = 1d0
  real(8) :: cdt = 0.9d0, cdtv = 0.25d0, dt = 1d-3
  integer :: precision = 0, kernel = 0
  character(len=512) :: infile = 'f_in.dat', outfile = 'f_out.dat'
  namelist /pcmx_case/ nx, ny, nz, nt, lmagnetic, lentropy, ldt, Lx, Ly, Lz, nu, eta, chi, cs0, gamma, &
                       cp, rho0, mu0, cdt, cdtv, dt, precision, kernel, infile, outfile

  real(8), parameter :: alpha_ts(3) = (/ 0d0, -5d0/9d0, -153d0/128d0 /)
  real(8), parameter :: beta_ts(3) = (/ 1d0/3d0, 15d0/16d0, 8d0/15d0 /)
  real(8), allocatable :: f(:,:,:,:)
  real(8) :: t, ds, red(PCMX_N_REDUCED)
  type(pcmx_params) :: par
  character(len=512) :: nmlfile
  integer :: mx, my, mz, mvar, it, itsub, u

  if (command_argument_count() < 1) stop 'usage: pcmx_run case.nml'
  call get_command_argument(1, nmlfile)
  open (newunit=u, file=trim(nmlfile), status='old', action='read')
  read (u, nml=pcmx_case)
  close (u)

  mvar = 4
  if (lentropy) mvar = mvar + 1
  if (lmagnetic) mvar = mvar + 3
  mx = nx + 6; my = ny + 6; mz = nz + 6
  allocate (f(mx, my, mz, mvar))
  open (newunit=u, file=trim(infile), access='stream', form='unformatted', status='old', action='read')
  read (u) f
  close (u)

  par%nx = nx; par%ny = ny; par%nz = nz; par%mvar = mvar
  par%iux = 0; par%ilnrho = 3
  par%iss = merge(4, -1, lentropy)
  par%iax = -1
  if (lmagnetic) par%iax = merge(5, 4, lentropy)
  par%dx = Lx/nx; par%dy = Ly/ny; par%dz = Lz/nz
  par%nu = nu; par%eta = eta; par%chi = chi
  par%cs0 = cs0; par%gamma = gamma; par%cp = cp; par%rho0 = rho0; par%mu0 = mu0
  par%cdt = cdt; par%cdtv = cdtv
  par%ldt = merge(1, 0, ldt); par%precision = precision; par%kernel = kernel
  par%lupw_uu = 0; par%lupw_lnrho = 0
  par%gravz = 0.; par%lperi_z = 1; par%bcz_bot = 0; par%bcz_top = 0; par%lupw_ss = 0
  par%omega = 0.; par%theta = 0.; par%lheatc_kprof = 0; par%cooling_profile = 0
  par%hcond0 = 0.; par%hcond1 = 0.; par%hcond2 = 0.; par%widthss = 0.; par%z1 = 0.; par%z2 = 0.
  par%luminosity = 0.; par%wheat = 0.; par%cool = 0.; par%wcool = 0.; par%cs2cool = 0.
  par%fbot = 0.; par%ftop = 0.; par%cs2bot = 0.; par%cs2top = 0.
  par%z0 = -Lz/2; par%lxyz = (/ Lx, Ly, Lz /)

  t = 0d0
  call check(pcmx_register(), 'register')
  call check(pcmx_initialize(f, mvar, par, t, 1), 'initialize')
  do it = 1, nt
    ds = 0d0
    do itsub = 1, 3
      call check(pcmx_before_boundary(0, itsub, t), 'before_boundary')
      call check(pcmx_rhs(itsub, t, dt), 'rhs')
      if (itsub == 1) then
        call check(pcmx_get_reduced_vars(red), 'get_reduced_vars')
        ds = 0d0
      else
        ds = alpha_ts(itsub)*ds
      endif
      ds = ds + 1d0
      t = t + ds*(dt*beta_ts(itsub))
      call check(pcmx_update_after_substep(), 'update_after_substep')
    enddo
    call check(pcmx_split_update(), 'split_update')
    write (*, '(i6,5es25.16)') it, t, dt, red(1), red(2), red(3)
  enddo
  call check(pcmx_copy_farray(f, mvar), 'copy_farray')
  open (newunit=u, file=trim(outfile), access='stream', form='unformatted', status='replace', action='write')
  write (u) f
  close (u)
  call check(pcmx_finalize(), 'finalize')

contains

  subroutine check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc < 0) then
      write (*, '(a)') 'pcmx_run: '//what//' failed: '//trim(pcmx_error_message())
      error stop 1
    endif
  end subroutine check

end program pcmx_run
