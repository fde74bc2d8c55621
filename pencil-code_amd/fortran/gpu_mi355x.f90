! $Id$
!
! MODULE_DOC: GPU module backed by libpcmx, a hand-written HIP engine for the
! MODULE_DOC: MI355X (gfx950): fused 6th-order RHS + 2N-RK3 substep kernels.
!
! CPARAM logical, parameter :: lgpu = .true.
!
!**************************************************************************
!
!  Drop-in replacement of gpu_astaroth.f90 (select with GPU = gpu_mi355x in
!  Makefile.local).  It exports the public list of src/gpu.h and talks to
!  the device only through the C ABI of include/pcmx.h, bound with
!  iso_c_binding in module pcmx_iface.  Supported physics (checked at
!  initialize_GPU where the flags are visible): hydro + log density +
!  (noentropy | entropy) + optional magnetic, nu-const viscosity, eta-const
!  resistivity, chi-const heat conduction, periodic Cartesian equidistant
!  grid, one process, itorder=3.
!
module GPU
!
  use Cdata
  use General, only: keep_compiler_quiet
  use Messages
  use iso_c_binding
  use pcmx_iface
!
  implicit none
!
  include 'gpu.h'
!
  logical :: ltest_bcs=.false.    !PAR-DOC: compare device and host periodic ghost zones at start
  logical :: ltest_rhs=.false.    !PAR-DOC: run the reference's test_rhs_gpu comparison (equ.f90)
  integer :: it_test_rhs=1
  integer :: gpu_precision=0      !PAR-DOC: 0 FP64, 1 mixed (FP32 stencils on box-mean offsets)
  integer :: gpu_kernel=0         !PAR-DOC: 0 auto, 1 direct, 2 z-marching LDS kernel, 3 split z-marching kernel
  logical :: lgpu_magnetic_forcing=.false. !PAR-DOC: the helical force acts on aa (set with forcing_run_pars lmagnetic_forcing=T)
  logical :: lgpu_tune=.false.    !PAR-DOC: time the split kernel's tile shapes and z-chunk lengths once at start
!
  namelist /gpu_run_pars/ ltest_bcs, ltest_rhs, it_test_rhs, gpu_precision, gpu_kernel, &
                          lgpu_magnetic_forcing, lgpu_tune
!
  real, dimension(n_gpu_reduced_vars) :: reduced_last = 0.
!
contains
!***********************************************************************
    subroutine pcmx_check(rc, caller)
!
!  Turns a negative libpcmx return code into fatal_error (the reference's
!  C++ side prints and exits instead, gpu_astaroth.cc:2074-2078).
!
      integer(c_int), intent(in) :: rc
      character(len=*), intent(in) :: caller
!
      if (rc < 0) call fatal_error(caller, trim(pcmx_error_message()))
!
    endsubroutine pcmx_check
!***********************************************************************
    subroutine register_GPU
!
      call pcmx_check(pcmx_register(), 'register_GPU')
!
    endsubroutine register_GPU
!***********************************************************************
    subroutine initialize_GPU(f)
!
!  Checks that the run uses only what the engine implements, passes the grid
!  and physics parameters explicitly (no generated headers), allocates the
!  device state and loads f (gpu_astaroth.f90:192-245).
!
      use SharedVariables, only: get_shared_variable
      use EquationOfState, only: rho0, eos_pushpars2c => pushpars2c
      use Gravity, only: gravity_pushpars2c => pushpars2c
      use Hydro, only: hydro_pushpars2c => pushpars2c
      use Density, only: density_pushpars2c => pushpars2c
      use Energy, only: energy_pushpars2c => pushpars2c
!
      real, dimension(:,:,:,:), intent(in) :: f
!
      type(pcmx_params) :: par
      real, pointer :: nu_p, eta_p, chi_p, cs20_p, gamma_p, cp_p
      real, dimension(:), pointer :: gz_p
      real(c_double), dimension(3) :: best
      integer(KIND=ikind8), dimension(500) :: p_par   ! the largest module list (entropy.f90:8672)
      logical, pointer :: lupw, lflag
      real, pointer :: rp
      integer, pointer :: ip
      character(len=512) :: str
      integer :: j, k
!
      str=''
      if (ncpus /= 1) str=trim(str)//', ncpus>1'
      if (.not. (lperi(1) .and. lperi(2))) str=trim(str)//', non-periodic x or y'
      if (.not. lcartesian_coords) str=trim(str)//', non-Cartesian coordinates'
      if (.not. all(lequidist)) str=trim(str)//', non-equidistant grid'
      if (.not. (lhydro .and. ldensity)) str=trim(str)//', needs hydro and density'
      if (ldensity_nolog) str=trim(str)//', linear density'
      if (ltemperature .or. lthermal_energy) str=trim(str)//', temperature/thermal_energy'
      if (itorder /= 3) str=trim(str)//', itorder/=3'
      if (ldt .and. (ddt > 0. .or. loutput_varn_at_exact_tsnap)) str=trim(str)//', dt growth limiter'
      if (str /= '') call fatal_error('initialize_GPU', &
          'gpu_mi355x does not implement: '//trim(str(3:)))
!
      call get_shared_variable('nu', nu_p)
      call get_shared_variable('cs20', cs20_p)
      call get_shared_variable('gamma', gamma_p)
      call get_shared_variable('cp', cp_p)
!
      par%nx = nx; par%ny = ny; par%nz = nz; par%mvar = mvar
      par%iux = iux - 1; par%ilnrho = ilnrho - 1
      par%iss = -1; par%iax = -1
      if (lentropy) par%iss = iss - 1
      if (lmagnetic) par%iax = iax - 1
      par%dx = dx; par%dy = dy; par%dz = dz
      par%nu = nu_p
      par%eta = 0.; par%chi = 0.
      if (lmagnetic) then
        call get_shared_variable('eta', eta_p)
        par%eta = eta_p
      endif
      if (lentropy) then
        call get_shared_variable('chi', chi_p)
        par%chi = chi_p
      endif
      par%cs0 = sqrt(cs20_p); par%gamma = gamma_p; par%cp = cp_p
      par%rho0 = rho0; par%mu0 = mu0
      par%cdt = cdt; par%cdtv = cdtv
      par%ldt = merge(1, 0, ldt)
      par%precision = gpu_precision
      par%kernel = gpu_kernel
!
!  Upwinding flags are private to Hydro and Density; their addresses reach
!  the GPU side through the modules' pushpars2c lists, as in the reference's
!  GPU path (hydro.f90:9367 p_par(2) = lupw_uu, density.f90:4593 p_par(5) =
!  lupw_lnrho).
!
      p_par = 0
      call hydro_pushpars2c(p_par)
      call c_f_pointer(transfer(p_par(2), c_null_ptr), lupw)
      par%lupw_uu = merge(1, 0, lupw)
      p_par = 0
      call density_pushpars2c(p_par)
      call c_f_pointer(transfer(p_par(5), c_null_ptr), lupw)
      par%lupw_lnrho = merge(1, 0, lupw)
      par%lupw_ss = 0
      if (lentropy) then                       ! entropy.f90:8704 p_par(29) = lupw_ss
        p_par = 0
        call energy_pushpars2c(p_par)
        call c_f_pointer(transfer(p_par(29), c_null_ptr), lupw)
        par%lupw_ss = merge(1, 0, lupw)
      endif
!
!  Constant vertical gravity (gravity_simple, gravz_profile='const'): the
!  profile the Gravity module shares (gravity_simple.f90:255, :516-523).
!
      par%gravz = 0.
      if (lgrav) then
        call get_shared_variable('gravz_zpencil', gz_p)
        if (any(gz_p /= gz_p(n1))) &
            call fatal_error('initialize_GPU', 'only gravz_profile=''const'' runs on the GPU')
        par%gravz = gz_p(n1)
      endif
!
!  z boundaries: periodic, or per variable 's', 'a', 'a2' below and above
!  (bcz12, split by the reference from run_pars bcz; bc_sym_z).
!
      par%lperi_z = merge(1, 0, lperi(3))
      par%bcz_bot = 0
      par%bcz_top = 0
      if (.not. lperi(3)) then
        do j = 1, mvar
          do k = 1, 2
            select case (trim(bcz12(j,k)))
            case ('s'); if (k == 1) then; par%bcz_bot(j) = PCMX_BC_S; else; par%bcz_top(j) = PCMX_BC_S; endif
            case ('a'); if (k == 1) then; par%bcz_bot(j) = PCMX_BC_A; else; par%bcz_top(j) = PCMX_BC_A; endif
            case ('a2'); if (k == 1) then; par%bcz_bot(j) = PCMX_BC_A2; else; par%bcz_top(j) = PCMX_BC_A2; endif
            case ('c1'); if (k == 1) then; par%bcz_bot(j) = PCMX_BC_C1; else; par%bcz_top(j) = PCMX_BC_C1; endif
            case ('cT'); if (k == 1) then; par%bcz_bot(j) = PCMX_BC_CT; else; par%bcz_top(j) = PCMX_BC_CT; endif
            case default
              call fatal_error('initialize_GPU', 'bcz='''//trim(bcz12(j,k))// &
                  ''' is not provided on the GPU (s, a, a2; c1, cT on ss)')
            endselect
          enddo
        enddo
      endif
!
!  The layer physics of samples/conv-slab.  Rotation: Omega and theta of
!  Cdata when the Coriolis force is on (hydro.f90:9380 p_par(15) =
!  lcoriolis_force, which initialize_hydro ands with lrotation).  K-profile
!  conduction, heating and cooling, and the constants of 'c1' / 'cT', through
!  the Energy (entropy.f90:8667-), EquationOfState (eos_idealgas.f90:4598-4599)
!  and Gravity (gravity_simple.f90:1603-1604) pushpars2c lists.
!
      par%omega = 0.; par%theta = 0.; par%lheatc_kprof = 0; par%cooling_profile = 0
      par%hcond0 = 0.; par%hcond1 = 0.; par%hcond2 = 0.; par%widthss = 0.; par%z1 = 0.; par%z2 = 0.
      par%luminosity = 0.; par%wheat = 0.; par%cool = 0.; par%wcool = 0.; par%cs2cool = 0.
      par%fbot = 0.; par%ftop = 0.; par%cs2bot = 0.; par%cs2top = 0.
      par%z0 = xyz0(3); par%lxyz = Lxyz
      p_par = 0
      call hydro_pushpars2c(p_par)
      call c_f_pointer(transfer(p_par(15), c_null_ptr), lflag)
      if (lflag .and. Omega /= 0.) then
        par%omega = Omega; par%theta = theta
      endif
      if (lentropy) then
        p_par = 0
        call energy_pushpars2c(p_par)
        call c_f_pointer(transfer(p_par(116), c_null_ptr), lflag)          ! lheatc_kprof
        if (lflag) then
          if (.not. lgrav) call fatal_error('initialize_GPU', 'K-profile conduction on the GPU needs gravity')
          par%lheatc_kprof = 1
          call c_f_pointer(transfer(p_par(72), c_null_ptr), rp); par%hcond0 = rp
          call c_f_pointer(transfer(p_par(139), c_null_ptr), lflag)        ! lmultilayer
          if (lflag) then
            call c_f_pointer(transfer(p_par(73), c_null_ptr), rp); par%hcond1 = rp
            call c_f_pointer(transfer(p_par(76), c_null_ptr), rp); par%hcond2 = rp
          else
            par%hcond1 = 1.; par%hcond2 = 1.                               ! hcond = hcond0
          endif
          call c_f_pointer(transfer(p_par(31), c_null_ptr), rp); par%widthss = rp
          call c_f_pointer(transfer(p_par(16), c_null_ptr), rp); par%fbot = rp
          call c_f_pointer(transfer(p_par(17), c_null_ptr), rp); par%ftop = rp
        endif
        call c_f_pointer(transfer(p_par(167), c_null_ptr), lflag)          ! lcalc_heat_cool
        if (lflag) then
          call c_f_pointer(transfer(p_par(133), c_null_ptr), lflag)        ! lcooling_to_cs2cool
          if (.not. lflag) call fatal_error('initialize_GPU', 'cooling on the GPU needs lcooling_to_cs2cool')
          call c_f_pointer(transfer(p_par(192), c_null_ptr), ip)           ! enum_cooling_profile
          if (ip == enum_gaussian_string) then
            par%cooling_profile = 0
          elseif (ip == enum_step_string) then
            par%cooling_profile = 1
          else
            call fatal_error('initialize_GPU', 'cooling_profile is not gaussian or step')
          endif
          call c_f_pointer(transfer(p_par(34), c_null_ptr), rp); par%luminosity = rp
          call c_f_pointer(transfer(p_par(35), c_null_ptr), rp); par%wheat = rp
          call c_f_pointer(transfer(p_par(36), c_null_ptr), rp); par%cool = rp
          call c_f_pointer(transfer(p_par(43), c_null_ptr), rp); par%wcool = rp
          call c_f_pointer(transfer(p_par(53), c_null_ptr), rp); par%cs2cool = rp
        endif
        p_par = 0
        call eos_pushpars2c(p_par)
        call c_f_pointer(transfer(p_par(10), c_null_ptr), rp); par%cs2bot = rp
        call c_f_pointer(transfer(p_par(11), c_null_ptr), rp); par%cs2top = rp
      endif
      if (lgrav .and. par%lheatc_kprof == 1) then
        p_par = 0
        call gravity_pushpars2c(p_par)
        call c_f_pointer(transfer(p_par(23), c_null_ptr), rp); par%z1 = rp
        call c_f_pointer(transfer(p_par(24), c_null_ptr), rp); par%z2 = rp
      endif
!
      call pcmx_check(pcmx_initialize(f, size(f,4), par, real(t, c_double), 0), 'initialize_GPU')
      if (nt > 0) call load_farray_to_GPU(f)
      if (ltest_bcs) call test_gpu_bcs
      if (lgpu_tune .and. nt > 0) then
!  (only the split kernel has plans to choose from; other kernels keep theirs)
        if (pcmx_tune(best) == 0) then
          if (lroot) print '(a,f8.3,a,i3,a,i5)', 'initialize_GPU: tuned split kernel: ', best(1), &
              ' ms per launch, tile width ', nint(best(2)), ', z chunk ', nint(best(3))
        elseif (lroot) then
          print '(a)', 'initialize_GPU: no tuning: '//trim(pcmx_error_message())
        endif
      endif
!
    endsubroutine initialize_GPU
!***********************************************************************
    subroutine read_gpu_run_pars(iomsg)
!
      use File_io, only: parallel_unit
!
      character(len=*), intent(out) :: iomsg
      integer :: iostat
!
      read(parallel_unit, NML=gpu_run_pars, IOSTAT=iostat, IOMSG=iomsg)
      if (iostat==0) iomsg=""
!
    endsubroutine read_gpu_run_pars
!***********************************************************************
    subroutine write_gpu_run_pars(unit)
!
      integer, intent(in) :: unit
!
      write(unit, NML=gpu_run_pars)
!
    endsubroutine write_gpu_run_pars
!***********************************************************************
    subroutine finalize_GPU
!
      call pcmx_check(pcmx_finalize(), 'finalize_GPU')
!
    endsubroutine finalize_GPU
!***********************************************************************
    subroutine get_farray_ptr_gpu
!
!  The device state is interior-only and stays private to the library.
!
    endsubroutine get_farray_ptr_gpu
!***********************************************************************
    subroutine rhs_GPU(f, isubstep)
!
!  One RK substep on the device: RHS, df = alpha*df + RHS, f += dt*beta*df
!  (gpu_astaroth.f90:285-295).  With a Courant-limited step the engine
!  computes dt1 from this step's first-substep RHS, before the update, as the
!  CPU path does (equ.f90:408); set_dt then forms dt = 1/dt1 exactly as the
!  device did.
!
      use Sub, only: set_dt
!
      real, dimension (mx,my,mz,mfarray), intent(inout) :: f
      integer, intent(in) :: isubstep
!
      real(c_double) :: dt_c
      real(c_double), dimension(PCMX_N_REDUCED) :: red
!
      call keep_compiler_quiet(f)
      if (lforcing .and. isubstep == itorder) call update_forcing_gpu
      dt_c = dt
      call pcmx_check(pcmx_rhs(isubstep, real(t, c_double), dt_c), 'rhs_GPU')
      if (isubstep == 1) then
        call pcmx_check(pcmx_get_reduced_vars(red), 'rhs_GPU')
        reduced_last = red
        if (ldt) call set_dt(real(red(4)))
      endif
!
    endsubroutine rhs_GPU
!***********************************************************************
    subroutine update_forcing_gpu
!
!  A new helical force for the step in progress, drawn by the reference's
!  own forcing_pars_hel (forcing.f90:1731-1746, with this step's dt) and
!  handed to the device, which adds it after the last substep; the
!  reference's GPU path does the same at its last substep
!  (astaroth/gpu_astaroth.cc:1589-1602, astaroth/forcing.h:13-29).  Only
!  iforce='helical' on its default path is provided.
!
      use Forcing, only: forcing_pars_hel, forcing_pushpars2c => pushpars2c
!
      integer(KIND=ikind8), dimension(500) :: p_par   ! the largest module list (entropy.f90:8672)
      logical, dimension(:), pointer :: lhel
      logical, pointer :: lsecond
      real, pointer :: k1_ff_p
      integer, pointer :: zsym
      real, dimension(3) :: coef1, coef2, coef3, fda, kk
      real :: phase, fact
      type(pcmx_forcing) :: frc
!
!  Flags and parameters through the module's pushpars2c list
!  (forcing.f90:6599-6712): k1_ff (1), iforcing_zsym (3), lforce_helical (82),
!  lsecond_force (83).
!
      p_par = 0
      call forcing_pushpars2c(p_par)
      call c_f_pointer(transfer(p_par(82), c_null_ptr), lhel, [2])
      call c_f_pointer(transfer(p_par(83), c_null_ptr), lsecond)
      call c_f_pointer(transfer(p_par(1), c_null_ptr), k1_ff_p)
      call c_f_pointer(transfer(p_par(3), c_null_ptr), zsym)
      if (.not. lhel(1)) call fatal_error('update_forcing_gpu', 'only iforce=''helical'' runs on the GPU')
      if (lsecond) call fatal_error('update_forcing_gpu', 'a second force is not provided on the GPU')
      if (zsym /= 0) call fatal_error('update_forcing_gpu', 'iforcing_zsym /= 0 is not provided on the GPU')
!
      call forcing_pars_hel(coef1, coef2, coef3, fda, kk, phase, fact)
      frc%active = 1
      frc%target = merge(iax - 1, iux - 1, lgpu_magnetic_forcing)
      frc%scale_sqrt_dt = 0
      frc%reserved = 0
      frc%kk = kk
      frc%phase = phase
      frc%fact = fact
      frc%coef1 = coef1
      frc%coef2 = coef2
      frc%fda = fda
      frc%k1_ff = k1_ff_p
      frc%xyz0 = xyz0
      call pcmx_check(pcmx_set_forcing(frc), 'update_forcing_gpu')
!
    endsubroutine update_forcing_gpu
!***********************************************************************
    subroutine before_boundary_gpu(f, lrmv, isubstep, t)
!
      real, dimension (mx,my,mz,mfarray), intent(inout) :: f
      logical :: lrmv
      integer, intent(in) :: isubstep
      real(KIND=rkind8), intent(in) :: t
!
      call keep_compiler_quiet(f)
      call pcmx_check(pcmx_before_boundary(merge(1, 0, lrmv), isubstep, real(t, c_double)), &
                      'before_boundary_gpu')
!
    endsubroutine before_boundary_gpu
!***********************************************************************
    subroutine update_after_substep_gpu
!
      call pcmx_check(pcmx_update_after_substep(), 'update_after_substep_gpu')
!
    endsubroutine update_after_substep_gpu
!***********************************************************************
    subroutine gpu_prepare_for_first_substep
!
!  Initial dt for Courant-limited runs (gpu_astaroth.cc:2456-2482).
!
      use Sub, only: set_dt
!
      real(c_double) :: dt1
!
      call pcmx_check(pcmx_prepare_for_first_substep(real(t, c_double), dt1), &
                      'gpu_prepare_for_first_substep')
      if (ldt) call set_dt(real(dt1))
!
    endsubroutine gpu_prepare_for_first_substep
!***********************************************************************
    function get_ptr_GPU(ind1, ind2, lout) result(pFarr)
!
      integer, optional :: ind1, ind2
      logical, optional :: lout
      real, dimension(:,:,:,:), pointer :: pFarr
!
      nullify(pFarr)
      call fatal_error('get_ptr_GPU', 'device arrays are not exposed by gpu_mi355x')
!
    endfunction get_ptr_GPU
!***********************************************************************
    function get_ptr_GPU_training(ind1, ind2, lout) result(pFarr)
!
      integer :: ind1
      integer, optional :: ind2
      logical, optional :: lout
      real, dimension(:,:,:,:,:), pointer :: pFarr
!
      nullify(pFarr)
      call fatal_error('get_ptr_GPU_training', 'training is not available with gpu_mi355x')
!
    endfunction get_ptr_GPU_training
!***********************************************************************
    subroutine copy_farray_from_GPU(f, nowait_)
!
!  Device state -> host f, interior and periodic ghosts of the mvar evolved
!  variables (gpu_astaroth.f90:390-412).  The copy is synchronous.
!
      real, dimension (mx,my,mz,mfarray), intent(out) :: f
      logical, optional :: nowait_
!
      call pcmx_check(pcmx_copy_farray(f, mfarray), 'copy_farray_from_GPU')
!
    endsubroutine copy_farray_from_GPU
!***********************************************************************
    subroutine load_farray_to_GPU(f)
!
      real, dimension (mx,my,mz,mfarray), intent(in) :: f
!
      call pcmx_check(pcmx_load_farray(f, mfarray), 'load_farray_to_GPU')
!
    endsubroutine load_farray_to_GPU
!***********************************************************************
    subroutine reload_GPU_config
!
!  RELOAD: push the possibly changed run parameters (gpu_astaroth.f90:423-427).
!
      use SharedVariables, only: get_shared_variable
!
      real, pointer :: p
      integer(c_int) :: rc
!
      call get_shared_variable('nu', p); rc = pcmx_update_by_name('nu'//c_null_char, real(p, c_double))
      call pcmx_check(rc, 'reload_GPU_config')
      if (lmagnetic) then
        call get_shared_variable('eta', p); rc = pcmx_update_by_name('eta'//c_null_char, real(p, c_double))
        call pcmx_check(rc, 'reload_GPU_config')
      endif
      if (lentropy) then
        call get_shared_variable('chi', p); rc = pcmx_update_by_name('chi'//c_null_char, real(p, c_double))
        call pcmx_check(rc, 'reload_GPU_config')
      endif
      rc = pcmx_update_by_name('cdt'//c_null_char, real(cdt, c_double)); call pcmx_check(rc, 'reload_GPU_config')
      rc = pcmx_update_by_name('cdtv'//c_null_char, real(cdtv, c_double)); call pcmx_check(rc, 'reload_GPU_config')
!
    endsubroutine reload_GPU_config
!***********************************************************************
    subroutine update_on_gpu_vec(index, varname, value)
!
      integer, intent(inout) :: index
      character(len=*), optional :: varname
      real, dimension(3), optional :: value
!
      call fatal_error('update_on_gpu_vec', 'no vector parameters on this path: '//trim(varname))
!
    endsubroutine update_on_gpu_vec
!***********************************************************************
    subroutine update_on_gpu(index, varname, value)
!
!  index<0: look the parameter up by name and return its index; an unknown
!  name is fatal, as in gpu_astaroth.f90:450-473.
!
      integer, intent(inout) :: index
      character(len=*), optional :: varname
      real, optional :: value
!
      integer(c_int) :: rc
!
      if (.not. present(value)) return
      if (index >= 0) then
        call pcmx_check(pcmx_update_by_index(int(index, c_int), real(value, c_double)), 'update_on_gpu')
      else
        rc = pcmx_update_by_name(trim(varname)//c_null_char, real(value, c_double))
        if (rc < 0) call fatal_error('update_on_gpu', 'variable "'//trim(varname)//'" not found')
        index = rc
      endif
!
    endsubroutine update_on_gpu
!***********************************************************************
    subroutine radtransfer_gpu
!
    endsubroutine radtransfer_gpu
!***********************************************************************
    subroutine get_gpu_reduced_vars(dst)
!
!  dst(1:5) = urms, umax, brms, dt1_max, dt of the last first substep;
!  dst(6:) = box means of the evolved variables.
!
      real, dimension(20) :: dst
      real(c_double), dimension(PCMX_N_REDUCED) :: red
!
      call pcmx_check(pcmx_get_reduced_vars(red), 'get_gpu_reduced_vars')
      dst = red
!
    endsubroutine get_gpu_reduced_vars
!***********************************************************************
    subroutine test_gpu_bcs
!
      real(c_double) :: maxdiff
!
      call pcmx_check(pcmx_test_bcs(maxdiff), 'test_gpu_bcs')
      if (lroot) print*, 'test_gpu_bcs: max |device - host| over ghost zones =', maxdiff
      if (maxdiff /= 0.) call fatal_error('test_gpu_bcs', 'device ghost zones differ from host')
!
    endsubroutine test_gpu_bcs
!***********************************************************************
    subroutine split_update_gpu(f)
!
      real, dimension (mx,my,mz,mfarray), intent(inout) :: f
!
      call keep_compiler_quiet(f)
      call pcmx_check(pcmx_split_update(), 'split_update_gpu')
!
    endsubroutine split_update_gpu
!***********************************************************************
    subroutine pushpars2c(p_par)
!
!  Parameters reach libpcmx through pcmx_params, not generated headers.
!
      integer, parameter :: n_pars=20
      integer(KIND=ikind8), dimension(n_pars) :: p_par
!
      call keep_compiler_quiet(p_par(1))
!
    endsubroutine pushpars2c
!***********************************************************************
    subroutine TF_train_gpu(model_name, loss, itsub, t)
      real :: loss
      integer :: itsub
      real(KIND=rkind8), intent(IN) :: t
      character(len=*), intent(in) :: model_name
      call fatal_error('TF_train_gpu', 'training is not available with gpu_mi355x')
    endsubroutine TF_train_gpu
!***********************************************************************
    subroutine TF_infer_gpu(model_name, flag)
      integer :: flag
      character(len=*), intent(in) :: model_name
      call fatal_error('TF_infer_gpu', 'training is not available with gpu_mi355x')
    endsubroutine TF_infer_gpu
!***********************************************************************
    subroutine TF_create_model(model_name, config_file_path, lmpicomm)
      logical :: lmpicomm
      character(len=*), intent(in) :: model_name, config_file_path
      call fatal_error('TF_create_model', 'training is not available with gpu_mi355x')
    endsubroutine TF_create_model
!***********************************************************************
    subroutine tau_snapshots()
    endsubroutine tau_snapshots
!***********************************************************************
    subroutine TF_load_model(model_name, fname)
      character(len=*), intent(in) :: model_name, fname
      call fatal_error('TF_load_model', 'training is not available with gpu_mi355x')
    endsubroutine TF_load_model
!***********************************************************************
    subroutine TF_load_model_checkpoint(model_name, checkpoint_dir)
      character(len=*), intent(in) :: model_name, checkpoint_dir
      call fatal_error('TF_load_model_checkpoint', 'training is not available with gpu_mi355x')
    endsubroutine TF_load_model_checkpoint
!***********************************************************************
    subroutine TF_save_model(model_name, fname)
      character(len=*), intent(in) :: model_name, fname
      call fatal_error('TF_save_model', 'training is not available with gpu_mi355x')
    endsubroutine TF_save_model
!***********************************************************************
    subroutine TF_save_checkpoint(model_name, checkpoint_dir)
      character(len=*), intent(in) :: model_name, checkpoint_dir
      call fatal_error('TF_save_checkpoint', 'training is not available with gpu_mi355x')
    endsubroutine TF_save_checkpoint
!***********************************************************************
endmodule GPU
