/*
 * pcmx.h -- C ABI of the MI355X RHS + RK3 engine (libpcmx.so).
 *
 * This is the drop-in boundary that a Pencil Code `module GPU`
 * (GPU = gpu_mi355x in Makefile.local) binds through iso_c_binding.  Each
 * entry point replaces one C function of the reference's Astaroth plugin
 * (src/gpu_astaroth_ansi.c, called from src/gpu_astaroth.f90); the citation
 * is given at each declaration.  Arguments are plain C: ints, doubles and
 * pointers, no device or torch types.  Every function returns PCMX_OK (0)
 * or a negative PCMX_ERR_* code; pcmx_last_error() then holds a message
 * (the reference prints and exits instead, gpu_astaroth.cc:2074-2078; the
 * Fortran side turns a negative code into fatal_error).
 *
 * State layout on the host (the reference f-array, src/cparam.f90:57-64):
 *   f(mx,my,mz,mfarray) real(8), column-major, x fastest, mx = nx+6 etc.
 *   variable slot v (0-based) starts at f + v*mx*my*mz.
 * Only the mvar evolved variables are transferred; slots are given by
 * pcmx_params (registration order, src/register.f90:137-146).
 *
 * Threading: all calls from one host thread (the OpenMP master, as in the
 * reference); device work is ordered on one library-owned HIP stream.
 */
#ifndef PCMX_H
#define PCMX_H

#ifdef __cplusplus
extern "C" {
#endif

#define PCMX_OK 0
#define PCMX_ERR_ARG (-1)      /* invalid argument or unsupported configuration */
#define PCMX_ERR_STATE (-2)    /* called before pcmx_initialize / after finalize */
#define PCMX_ERR_HIP (-3)      /* HIP runtime error (message names the call) */
#define PCMX_ERR_NOTFOUND (-4) /* update by name: unknown parameter */

#define PCMX_PREC_FP64 0  /* stencils, RHS and state in double */
#define PCMX_PREC_MIXED 1 /* stencils + RHS in float on offsets from the box
                             mean; state, RK accumulation, dt and diagnostics
                             in double */

#define PCMX_N_REDUCED 20 /* n_gpu_reduced_vars, src/gpu.h:12 */

/* Run parameters.  Replaces the generated-header parameter export of the
 * reference (pushpars2c / copy_addr, gpu_astaroth.f90:503-518) with an
 * explicit argument list. */
typedef struct pcmx_params {
  int nx, ny, nz;            /* interior grid (cparam.local nxgrid...; one process) */
  int mvar;                  /* evolved variables: 4 hydro, 5 +entropy, 7 MHD, 8 MHD+entropy */
  int iux, ilnrho, iss, iax; /* 0-based slots; iss / iax = -1 when absent */
  double dx, dy, dz;         /* grid spacing (grid.f90:123-125) */
  double nu;                 /* viscosity_run_pars nu, ivisc='nu-const' */
  double eta;                /* magnetic_run_pars eta, iresistivity='eta-const' */
  double chi;                /* entropy_run_pars chi, iheatcond='chi-const' */
  double cs0, gamma, cp, rho0, mu0; /* eos_init_pars / cdata */
  double cdt, cdtv;          /* run_pars Courant factors (cdata.f90:235-239) */
  int ldt;                   /* 1: Courant-limited dt (run_pars dt=0, timestep.f90:67) */
  int precision;             /* PCMX_PREC_* */
  int kernel;                /* 0: auto, 1: direct (one thread per cell), 2: z-marching LDS,
                                3: z-marching LDS with flow/field halves on separate waves */
  int lupw_uu;               /* hydro_run_pars lupw_uu: upwinded u.grad u (hydro.f90:3662,
                                sub.f90:7541-7574); the split kernel takes it together with
                                lupw_lnrho and without entropy, the direct kernel in any case */
  int lupw_lnrho;            /* density_run_pars lupw_lnrho (density.f90:2772-2773) */
  double gravz;              /* grav_run_pars gravz, gravz_profile='const' (gravity_simple.f90:
                                addgravity :1083-1162): df(uz) += gravz after the magnetic terms */
  int lperi_z;               /* init_pars lperi(3): 1 periodic z; 0 the bcz rules below, with
                                boundary points at the first and last z plane (dz = Lz/(nz-1)) */
  int bcz_bot[8], bcz_top[8];/* run_pars bcz per variable slot when lperi_z == 0: PCMX_BC_* */
  int lupw_ss;               /* entropy_run_pars lupw_ss: upwinded u.grad ss (entropy.f90:3576);
                                entropy runs, direct kernel */
  /* Layer physics of samples/conv-slab: on the split kernel (kernel 3) with entropy,
   * lupw_lnrho = lupw_ss = 1 and 32x8 tiles, else on the direct kernel (kernel 1);
   * kernel 0 chooses.  All zero: off. */
  double omega, theta;       /* hydro_run_pars Omega, theta (degrees): Coriolis force
                                (coriolis_cartesian, hydro.f90:6339-6413) */
  int lheatc_kprof;          /* entropy_run_pars iheatcond='K-profile' (lmultilayer, lgravz):
                                ds/dt += K(z)/rho (del2 lnT + grad lnT.(grad lnT + grad ln K)) */
  int cooling_profile;       /* 0 'gaussian' about the top, 1 'step' at z2 (entropy.f90:1319-1368) */
  double hcond0, hcond1, hcond2, widthss; /* K(z) = hcond0 (1 + (hcond1-1) step(z,z1,-widthss)
                                             + (hcond2-1) step(z,z2,widthss)) (entropy.f90:7741-7769) */
  double z1, z2;             /* grav_run_pars z1, z2 */
  double luminosity, wheat;  /* bottom heating, a half-Gaussian of width wheat (:1374-1386) */
  double cool, wcool, cs2cool; /* top cooling towards cs2cool (get_heat_cool_gravz :6761-6828);
                                  luminosity or cool needs cs2cool > 0 */
  double fbot, ftop;         /* heat fluxes of PCMX_BC_C1 (initialize_energy :845-890) */
  double cs2bot, cs2top;     /* sound speeds squared of PCMX_BC_CT */
  double z0;                 /* init_pars xyz0(3): z of the first grid point (non-periodic z) or of
                                the lower box face (periodic z) */
  double lxyz[3];            /* init_pars Lxyz (heating normalisation, top of the box) */
} pcmx_params;

/* z boundary conditions of bc_sym_z (boundcond.f90:3842-3900) */
#define PCMX_BC_S 1  /* 's':  f(n1-i) = f(n1+i) */
#define PCMX_BC_A 2  /* 'a':  f(n1-i) = -f(n1+i), f(n1) = 0 */
#define PCMX_BC_A2 3 /* 'a2': f(n1-i) = 2 f(n1) - f(n1+i) */
/* entropy-only rules (slot iss), with lnrho's rule applied first */
#define PCMX_BC_C1 4 /* 'c1': constant heat flux fbot / ftop through K(z) at the boundary
                        (bc_ss_flux, entropy_bcs.f90:93-235; needs lheatc_kprof) */
#define PCMX_BC_CT 5 /* 'cT': constant temperature, cs2 = cs2bot / cs2top at the boundary
                        (bc_ss_temp_z, entropy_bcs.f90:1329-1457) */

/* Last error message (thread-local to the calling thread). */
const char* pcmx_last_error(void);

/* register_GPU -> register_gpu_c (gpu_astaroth_ansi.c:125-130).  No-op. */
int pcmx_register(void);

/* initialize_GPU -> initialize_gpu_c(f, comm, t, nt, ...)
 * (gpu_astaroth.f90:192-245, gpu_astaroth_ansi.c:110-124, gpu_astaroth.cc:2009-2122).
 * Allocates device state for the grid and physics in *par.  When nt > 0 the
 * host f (layout above, dims mx,my,mz,mfarray) is loaded, as initialize_GPU
 * does through load_farray_to_GPU.  f may be NULL when nt == 0. */
int pcmx_initialize(const double* f, int mfarray, const pcmx_params* par, double t, int nt);

/* finalize_GPU -> finalize_gpu_c (gpu_astaroth_ansi.c:131-136, gpu_astaroth.cc:2243-2272). */
int pcmx_finalize(void);

/* load_farray_to_GPU -> load_farray_c (gpu_astaroth.f90:414-421, gpu_astaroth.cc:1947-1971).
 * Copies the mvar evolved variables of the host f into device state; ghost
 * zones of f are ignored (the device keeps the periodic interior only). */
int pcmx_load_farray(const double* f, int mfarray);

/* copy_farray_from_GPU -> copy_farray_c (gpu_astaroth.f90:390-412, gpu_astaroth.cc:1719-1814).
 * Writes the device state into the host f, interior and periodic ghost zones
 * of the mvar evolved variables. */
int pcmx_copy_farray(double* f, int mfarray);

/* before_boundary_gpu -> before_boundary_gpu_c (gpu_astaroth.f90:297-313,
 * gpu_astaroth.cc:1320-1353).  Periodic ghosts are formed inside the RHS
 * kernel, so nothing is done; kept for the call sequence of equ.f90:198. */
int pcmx_before_boundary(int lrmv, int isubstep, double t);

/* rhs_GPU -> rhs_gpu_c(isubstep, t) -> substepGPU (gpu_astaroth.f90:285-295,
 * gpu_astaroth_ansi.c:170-208, gpu_astaroth.cc:1672-1717).
 * Runs RK substep isubstep (1..3): RHS of every evolved variable, the 2N
 * accumulation df = alpha*df + RHS and the update f += dt*beta*df
 * (timestep.f90:110-161).  *dt is the step; when par->ldt is set and
 * isubstep == 1, the library computes dt = 1/max(dt1) from the state before
 * the update (equ.f90:408, sub.f90:4223-4254) and returns it in *dt.
 * On isubstep 1 the urms/umax/brms diagnostics of the pre-update state are
 * reduced on the device (readable with pcmx_get_reduced_vars). */
int pcmx_rhs(int isubstep, double t, double* dt);

/* update_after_substep_gpu -> update_after_substep_gpu_c (gpu_astaroth.cc:1515-1533).  No-op. */
int pcmx_update_after_substep(void);

/* gpu_prepare_for_first_substep -> gpu_prepare_for_first_substep_c
 * (gpu_astaroth.f90:321-325, gpu_astaroth.cc:2456-2482).  Computes dt1_max of
 * the current state without advancing it; returned in *dt1 (may be NULL). */
int pcmx_prepare_for_first_substep(double t, double* dt1);

/* get_gpu_reduced_vars -> get_gpu_reduced_vars_c (gpu_astaroth.f90:481-487,
 * gpu_astaroth.cc:2486-2530).  dst[PCMX_N_REDUCED] receives, from the last
 * substep 1: 0 urms, 1 umax, 2 brms, 3 dt1_max, 4 dt, 5.. box means of the
 * mvar variables; unused slots are zero. */
int pcmx_get_reduced_vars(double* dst);

/* update_on_gpu(index,varname,value) -> update_on_gpu_scal_by_name_c /
 * _by_ind_c (gpu_astaroth.f90:450-473, gpu_astaroth_ansi.c:236-259).
 * Names: nu, eta, chi, cs0, cdt, cdtv.  By name returns the index (>= 0) or
 * PCMX_ERR_NOTFOUND, like updateInConfigScalName. */
int pcmx_update_by_name(const char* varname, double value);
int pcmx_update_by_index(int index, double value);

/* test_gpu_bcs -> test_bcs_c (gpu_astaroth.f90:489-493, gpu_astaroth.cc:2283-2454).
 * Fills a random state, forms the periodic ghost zones on the device and on
 * the host and returns the largest difference in *maxdiff (0 when equal). */
int pcmx_test_bcs(double* maxdiff);

/* split_update_gpu -> split_update_gpu_c (gpu_astaroth.f90:495-501).  No-op
 * (no implicit / split physics on this path). */
int pcmx_split_update(void);

/* Advance nsteps full RK3 steps on the device with no host synchronisation
 * between substeps (the loop of timestep.f90:206-273 without the host
 * between substeps); *t and *dt are updated as the reference advances them.
 * With ldiag, urms/umax/brms are reduced on every step's first substep. */
int pcmx_time_step(int nsteps, double* t, double* dt, int ldiag);

/* Times the tile shapes and z-chunk lengths the grid allows for the split
 * kernel (kernel 3) on the current state, without changing it (the three
 * stages of one step, as pcmx_time_step runs them for fixed or Courant dt,
 * each reading the state and writing only the spare buffer and the 2N
 * registers), and keeps the fastest.  Call it between steps.
 * best[0] = mean ms per launch of the chosen plan, best[1] = tile width,
 * best[2] = z-chunk length (best may be NULL).  No counterpart in the
 * reference (its GPU library fixes its launch shapes at build time). */
int pcmx_tune(double* best);

/* Launch statistics of the last pcmx_time_step: averages in milliseconds of
 * the fused RHS kernel per substep, measured with HIP events on the library
 * stream.  out[0] = mean kernel ms, out[1] = number of kernels timed. */
int pcmx_kernel_timing(double* out);
int pcmx_set_timing(int enable);

/* ------------------------------------------------------------------------
 * Helical forcing (forcing_run_pars iforce='helical').
 *
 * pcmx_set_forcing replaces the reference GPU path's forcing update
 * (src/astaroth/forcing.h:13-29: forcing_pars_hel on the host, parameters
 * loaded as device uniforms) and its forcing term (src/astaroth/DSL/forcing/
 * pcstyleforcing.h:15-48, added at the last substep, DSL/solve.ac:64-67);
 * the CPU path adds the same term after each step (run.f90:575, addforce ->
 * forcing_hel, forcing.f90:2066-2474).  Default path only: no profiles,
 * omega_ff=0, force1_scl=1, iforcing_zsym=0, no discrete phases, one force.
 * After each full step (the end of substep 3 of pcmx_rhs / pcmx_time_step)
 * the library adds, for j = 0..2,
 *   f[target+j] += fda[j] * Re((coef1[j] + i coef2[j]) * fx(x) fy(y) fz(z)),
 *   fx = fact' exp(i (kk[0] k1_ff x + phase)), fy = exp(i kk[1] k1_ff y),
 *   fz = exp(i kk[2] k1_ff z)                    (fxyz_coefs_hel, :2003-2064)
 * at the cell centres x = xyz0[0] + (l + 1/2) dx (l = 0..nx-1), where
 * fact' = fact, or fact*sqrt(dt) of that step when scale_sqrt_dt != 0
 * (fconst_coefs_hel, :1773-2001, forms fact = force/ffnorm*sqrt(dt); pass
 * force/ffnorm to have dt applied by the library, as with Courant steps run
 * by pcmx_time_step).  The parameters stay in force until the next call;
 * active = 0 switches the forcing off.  target: 0-based slot of the forced
 * vector's x component, iux (hydro forcing) or iax (lmagnetic_forcing=T). */
typedef struct pcmx_forcing {
  int active;
  int target;
  int scale_sqrt_dt;
  int reserved;      /* 0 */
  double kk[3];      /* wave vector as forcing_pars_hel returns it */
  double phase, fact;
  double coef1[3], coef2[3], fda[3];
  double k1_ff;      /* forcing_run_pars k1_ff */
  double xyz0[3];    /* box origin (init_pars xyz0) */
} pcmx_forcing;
int pcmx_set_forcing(const pcmx_forcing* frc);

/* ------------------------------------------------------------------------
 * Snapshots written from the device.  No counterpart in the reference's
 * module GPU, which copies the whole f-array back (copy_farray_from_GPU,
 * equ.f90:307-317) before the host writes it (output_snap,
 * src/io_dist.f90:108-199).
 *
 * pcmx_write_snapshot writes the evolved variables in the reference's
 * var.dat layout for one process, without the persistent-data trailer
 * (lpersist=F):
 *   record 1: f(mx,my,mz,1:mvar) real(8) with periodic ghost zones
 *             (io_dist.f90:187);
 *   record 2: t, x(1:mx), y(1:my), z(1:mz), dx, dy, dz, real(8)
 *             (io_dist.f90:195);
 * x(i) = xyz0[0] + (i - 3.5)*dx for 1-based i: the periodic equidistant grid
 * of grid.f90:110-216 (the reference computes the same values through its
 * grid function, so they may differ in the last bits).  The state goes to
 * the file one variable at a time through two pinned 64 MiB chunks; the
 * host f-array is not touched.  Records follow the gfortran runtime's
 * unformatted sequential convention (4-byte byte counts around each record;
 * records over 2147483639 bytes split into subrecords whose leading count is
 * negated when another follows and whose trailing count is negated when one
 * precedes), which the reference's input_snap reads back. */
int pcmx_write_snapshot(const char* path, double t, const double* xyz0);

/* One record of that convention holding nparts arrays of doubles back to
 * back, written to path (appended when append != 0).  max_subrecord: the
 * subrecord length in bytes, 0 for the runtime's 2147483639.  A host-side
 * helper (time series, grid files, tests of the record layout); needs no
 * GPU and no pcmx_initialize. */
int pcmx_write_record(const char* path, int append, const double* const* parts, const long* counts, int nparts,
                      long max_subrecord);

/* ------------------------------------------------------------------------
 * Ensembles of independent boxes.  No counterpart in the reference's module
 * GPU (one box per run there); for callers that advance many small periodic
 * boxes, e.g. parameter scans.  All boxes share the physics of the params
 * passed to pcmx_ens_create (variables, EOS, precision: isothermal hydro or
 * MHD, fixed steps); each box has its own grid, spacing, viscosity,
 * resistivity and time step.  States are exchanged interior-only, as
 * f[var][z][y][x] doubles (C order), var in registration order.  One
 * ensemble per process, independent of the single-box engine above. */
typedef struct pcmx_box {
  int nx, ny, nz;      /* interior grid; nx, ny multiples of 16 (or of 32 and 8) */
  double dx, dy, dz;   /* grid spacing */
  double nu, eta;      /* viscosity, resistivity */
  double dt;           /* fixed time step */
} pcmx_box;

#define PCMX_ENS_NO_GRAPH 1 /* flags: advance with plain launches instead of a HIP graph */

int pcmx_ens_create(const pcmx_params* par, const pcmx_box* boxes, int nbox, int flags);
int pcmx_ens_destroy(void);
/* Load / copy back the interior state of box ibox (mvar*nz*ny*nx doubles). */
int pcmx_ens_load(int ibox, const double* f);
int pcmx_ens_copy(int ibox, double* f);
/* Advance every box nsteps RK3 steps (one persistent launch per substep;
 * pairs of steps replayed from a captured HIP graph).  With ldiag, urms,
 * umax, brms of each box are reduced on every step's first substep. */
int pcmx_ens_step(int nsteps, int ldiag);
/* dst[3*ibox + 0..2] = urms, umax, brms of the last diagnostic substep. */
int pcmx_ens_reduced(double* dst);
/* Times candidate tile shapes, z-chunk lengths and blocks per CU on the
 * loaded state (nsteps_probe steps each, the state restored afterwards) and
 * keeps the fastest. */
int pcmx_ens_tune(int nsteps_probe);
int pcmx_ens_set_timing(int enable);
/* out[0] mean ms per substep of the last timed pcmx_ens_step, out[1] substeps,
 * out[2..5] tile width, z-chunk limit, blocks per CU, work items. */
int pcmx_ens_timing(double* out);

#ifdef __cplusplus
}
#endif
#endif
