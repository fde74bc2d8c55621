/*
 * pc_oracle.h -- CPU restatement of the Pencil Code RHS + 2N-RK3 hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness oracle and the
 * CPU baseline; it is never linked into, loaded by, or called from the
 * product library (libpcmx).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may use it, and only as the checker.
 *
 * It follows the reference's CPU pencil loop function by function (file:line
 * citations are given at each routine in pc_oracle.c):
 *   src/deriv.f90      der_main, der2_main, derij_main (bidiagonal)
 *   src/sub.f90        gij, div_mn, curl_mn, traceless_strain, multmv_mn,
 *                      dot_mn, dot2_mn, cross_mn, multm2_sym_mn, del2v_etc,
 *                      gij_etc, set_dt
 *   src/hydro.f90, density.f90, viscosity.f90, noentropy.f90, entropy.f90,
 *   eos_idealgas.f90, magnetic.f90  (pencils and d*_dt RHS terms)
 *   src/equ.f90        rhs_cpu pencil loop, calc_maxadvec, set_dt1_max
 *   src/timestep.f90   advance_substeps (Williamson 2N-RK3)
 *   src/boundcond.f90  bc_per_x/y/z (periodic ghost fill)
 *   src/diagnostics.f90 urms / umax / brms normalisation
 *
 * Memory layout is the reference f-array: f(mx,my,mz,nvar), column-major,
 * x fastest, mx = nx + 2*NGHOST.  Variable slots follow registration order
 * (src/register.f90:137-146): uu = 0..2, lnrho = 3, ss = 4 if entropy,
 * aa = 4..6 (isothermal) or 5..7 (entropy).  Indices here are 0-based.
 *
 * Parity status of each function is recorded in the module header of
 * pc_oracle.c.
 */
#ifndef PC_ORACLE_H
#define PC_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define PCO_NGHOST 3

typedef struct {
  int nx, ny, nz;          /* interior grid */
  int lmagnetic;           /* MAGNETIC=magnetic */
  int lentropy;            /* ENTROPY=entropy (else noentropy) */
  double Lx, Ly, Lz;       /* box size (default 2*pi, start.f90 xyz0=-pi) */
  double nu, eta, chi;     /* nu-const viscosity, eta-const resistivity, chi-const conduction */
  double cs0, gamma, cp, rho0, mu0;
  double cdt, cdtv;        /* Courant factors (cdata.f90:235-239) */
  double dt;               /* fixed dt; 0 => Courant-limited (timestep.f90:67) */
  int lupw_uu;             /* hydro_run_pars lupw_uu: upwinded u.grad u (hydro.f90:3662) */
  int lupw_lnrho;          /* density_run_pars lupw_lnrho: upwinded u.grad lnrho (density.f90:2772-2773) */
  double gravz;            /* grav_run_pars gravz with gravz_profile='const' (gravity_simple.f90) */
  int lperi_z;             /* init_pars lperi(3); 0: the bcz rules below hold in z */
  int bcz_bot[8], bcz_top[8]; /* per variable: PCO_BC_S, PCO_BC_A, PCO_BC_A2 (run_pars bcz) */
  int lupw_ss;             /* entropy_run_pars lupw_ss: upwinded u.grad ss (entropy.f90:3576) */
  /* samples/conv-slab physics (lgravz):                                      */
  double omega, theta;     /* hydro_run_pars Omega, theta (degrees): Coriolis force
                              (coriolis_cartesian, hydro.f90:6339-6413) */
  int lheatc_kprof;        /* entropy_run_pars iheatcond='K-profile' with lmultilayer */
  int cooling_profile;     /* 0 'gaussian', 1 'step' (entropy.f90:1319-1368) */
  double hcond0, hcond1, hcond2, widthss; /* K(z) = hcond0*(1 + (hcond1-1) step(z,z1,-widthss)
                                             + (hcond2-1) step(z,z2,widthss)) */
  double z1, z2;           /* grav_run_pars z1, z2 */
  double luminosity, wheat;/* bottom heating (entropy.f90:1374-1386) */
  double cool, wcool, cs2cool; /* top cooling towards cs2cool (get_heat_cool_gravz) */
  double fbot, ftop;       /* heat fluxes of 'c1' (initialize_energy, entropy.f90:845-890) */
  double cs2bot, cs2top;   /* sound speed squared of 'cT' */
  double z0;               /* init_pars xyz0(3): z of the first grid point (non-periodic z) */
} pco_params;

#define PCO_BC_S 1   /* 's':  symmetric about the boundary point */
#define PCO_BC_A 2   /* 'a':  antisymmetric, boundary value 0 */
#define PCO_BC_A2 3  /* 'a2': antisymmetric about the boundary value */
#define PCO_BC_C1 4  /* 'c1' on ss: constant heat flux (bc_ss_flux, entropy_bcs.f90:93-235) */
#define PCO_BC_CT 5  /* 'cT' on ss: constant temperature (bc_ss_temp_z, entropy_bcs.f90:1329-1457) */

/* z profiles of the conv-slab physics at interior plane n (0-based):
 * out = {hcond, dlnhcond/dz, profz_heat, profz_cool} (get_gravz_heatcond,
 * entropy.f90:7741-7769; heating and cooling profiles :1319-1386). */
void pco_layer_profile(const pco_params* p, int n, double out[4]);

/* Helical forcing of one step, as forcing_pars_hel returns it
 * (forcing.f90:1707-2001) plus the grid origin and the forced variable. */
typedef struct {
  double kk[3], phase, fact;
  double coef1[3], coef2[3], fda[3];
  double k1_ff;
  double x0, y0, z0;       /* box origin (start.in xyz0) */
  int target;              /* first forced slot: 0 (uu) or iax (lmagnetic_forcing) */
} pco_forcing;

/* addforce with iforce='helical' on its default path (forcing.f90:1307-1395,
 * 2066-2474: unit profiles, omega_ff=0, dtforce=0, force1_scl=1) and
 * fxyz_coefs_hel (:2003-2064, iforcing_zsym=0, no discrete phases):
 * f(l,m,n,target+j) += fda(j)*real(cmplx(coef1(j),coef2(j))*fx(l)*fy(m)*fz(n)),
 * fx = fact*exp(i*(kk(1)*k1_ff*x+phase)), fy = exp(i*kk(2)*k1_ff*y),
 * fz = exp(i*kk(3)*k1_ff*z), x(l) = x0 + (l + 1/2)*dx on the interior. */
void pco_addforce(double* f, const pco_params* p, const pco_forcing* fo);

/* Variable slot helpers (0-based). */
int pco_nvar(const pco_params* p);
int pco_iss(const pco_params* p);   /* -1 when no entropy */
int pco_iax(const pco_params* p);   /* -1 when no magnetic */

/* Derived EOS constants as eos_idealgas.f90:initialize_eos computes them
 * (cs20, gamma_m1, gamma1, cp1, cv, cv1, lnrho0, lnTT0), in that order. */
void pco_eos_constants(const pco_params* p, double out[8]);

/* Periodic ghost fill of every variable, boundconds_x then _y then _z over
 * full extents (boundcond.f90:2387-2433, 2509-). */
void pco_fill_ghosts(double* f, int nx, int ny, int nz, int nvar);

/* boundconds_x/y/z (equ.f90:209-224) for the run's boundary conditions:
 * periodic in x and y, periodic or bcz in z (bc_sym_z, boundcond.f90:3842-3900;
 * 'a' sets the boundary value to 0).  Modifies f's ghost zones and, for 'a',
 * its boundary planes. */
void pco_boundconds(double* f, const pco_params* p);

/* Single stencil applications on a ghost-filled f for all interior points.
 * out has nx*ny*nz entries, x fastest.  k is the 0-based variable. */
void pco_der(const double* f, const pco_params* p, int k, int j, double* out);
void pco_der2(const double* f, const pco_params* p, int k, int j, double* out);
void pco_derij(const double* f, const pco_params* p, int k, int i, int j, double* out);

/* One pde() call: ghosts are filled, then for every pencil the module RHS
 * terms are added to df (which the caller pre-scales, as advance_substeps
 * does), in calc_all_rhs order.  When want_dt, dt1_max (the maximum of the
 * inverse Courant time step over the domain) is returned in *dt1.
 * When diag != NULL it receives {urms, umax, brms} of the input state.
 * nthreads <= 0 uses the OpenMP default. */
void pco_pde(double* f, double* df, const pco_params* p, int want_dt,
             double* dt1, double* diag, int nthreads);

/* nsteps full time steps (time_step / advance_substeps).  f and df are
 * full (mx,my,mz,nvar) arrays.  *dt is the fixed step, or on return the last
 * Courant step when p->dt == 0.  *t is advanced as the reference does.
 * diag_series (may be NULL) receives 3 values {urms, umax, brms} per step,
 * evaluated on the first substep (equ.f90:120). */
void pco_advance(double* f, double* df, const pco_params* p, int nsteps,
                 double* t, double* dt, double* diag_series, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
