/*
 * pc_oracle.c -- CPU restatement of the Pencil Code RHS + 2N-RK3 hot path.
 *
 * TEST INFRASTRUCTURE.  This file is the correctness oracle and the CPU
 * baseline (cpu_baseline.kind = "port").  It is not part of the product and
 * the product never calls it.
 *
 * The reference Fortran build cannot be produced in this container by the
 * rules of this project (it needs generated cparam.inc / pencil tables from
 * its own build system), so every function here is pinned instead by
 * closed forms, identities, linear theory and the reference's own printed
 * known answers; tests/test_oracle_*.py hold those checks.  Pin status:
 *   der/der2/derij ........ pinned (closed forms on sin waves, polynomial
 *                           exactness, bidiagonal structure)
 *   gij/curl/div/strain ... pinned (div curl = 0, curl curl = graddiv - del2,
 *                           S symmetric & traceless)
 *   hydro/density/visc .... pinned (linear sound wave, viscous shear decay,
 *                           Galilean/translation invariance)
 *   eos/entropy ........... pinned (reference-printed lnTT0, lnTT, cs2 at a
 *                           point; entropy-wave and heat-conduction decay)
 *   magnetic .............. pinned (resistive decay, linear Alfven wave,
 *                           J x B of a force-free field)
 *   set_dt1_max ........... pinned (closed-form dt1 of uniform states)
 *   RK3 ................... pinned (amplification factor 1+z+z^2/2+z^3/6)
 *   diagnostics ........... pinned (closed-form urms/brms of sine fields)
 *   nonlinear cross-term assembly at round-off: follows the reference's
 *   operation order; no reference-run vectors exist for it.
 *
 * Reference file:line citations are at each routine.  Indices are 0-based;
 * the Fortran's l1/m1/n1 are NGHOST here.
 */
#include "pc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NG PCO_NGHOST

/* Stencil coefficients, deriv.f90:36-40 (der2_type='standard') and
 * der_main's a=1/60 (deriv.f90:103). */
static const double DER_A = 1.0 / 60.0;
static const double DER2_C0 = -490.0 / 180.0;
static const double DER2_C1 = 270.0 / 180.0;
static const double DER2_C2 = -27.0 / 180.0;
static const double DER2_C3 = 2.0 / 180.0;

/* Williamson 2N-RK3, timestep.f90:43-44. */
static const double ALPHA_TS[3] = {0.0, -5.0 / 9.0, -153.0 / 128.0};
static const double BETA_TS[3] = {1.0 / 3.0, 15.0 / 16.0, 8.0 / 15.0};

int pco_nvar(const pco_params* p) { return 4 + (p->lentropy ? 1 : 0) + (p->lmagnetic ? 3 : 0); }
int pco_iss(const pco_params* p) { return p->lentropy ? 4 : -1; }
int pco_iax(const pco_params* p) { return p->lmagnetic ? (p->lentropy ? 5 : 4) : -1; }

/* eos_idealgas.f90:159-268 with rho0, cs0, cp, gamma given. */
void pco_eos_constants(const pco_params* p, double out[8]) {
  double gamma = p->gamma;
  double gamma_m1 = gamma - 1.0;
  double gamma1 = 1 / gamma;
  double lnrho0 = log(p->rho0);
  double cs20 = p->cs0 * p->cs0;
  double cp = p->cp;
  double cp1 = 1 / cp;
  double cv = gamma1 * cp;
  double cv1 = gamma * cp1;
  double lnTT0 = (gamma_m1 != 0.0) ? log(cs20 / (cp * gamma_m1)) : log(cs20 / cp);
  out[0] = cs20; out[1] = gamma_m1; out[2] = gamma1; out[3] = cp1;
  out[4] = cv; out[5] = cv1; out[6] = lnrho0; out[7] = lnTT0;
}

/* ------------------------------------------------------------------ */
/* grid geometry                                                       */
typedef struct {
  int nx, ny, nz, mx, my, mz;
  long mw;
  double dx_1, dy_1, dz_1, dxyz_2;
} geom;

static geom make_geom(const pco_params* p) {
  geom g;
  g.nx = p->nx; g.ny = p->ny; g.nz = p->nz;
  g.mx = p->nx + 2 * NG; g.my = p->ny + 2 * NG; g.mz = p->nz + 2 * NG;
  g.mw = (long)g.mx * g.my * g.mz;
  /* grid.f90:123-125 (periodic: dx = Lx/nxgrid, else Lx/(nxgrid-1)), dx_1 = 1/xprim (:495) */
  double dx = p->Lx / p->nx, dy = p->Ly / p->ny, dz = p->lperi_z ? p->Lz / p->nz : p->Lz / (p->nz - 1);
  g.dx_1 = 1.0 / dx; g.dy_1 = 1.0 / dy; g.dz_1 = 1.0 / dz;
  /* grid.f90:1063 (lmaximal_cdtv=F): dline_1(:,1)**2+dline_1(:,2)**2+dline_1(:,3)**2 */
  g.dxyz_2 = g.dx_1 * g.dx_1 + g.dy_1 * g.dy_1 + g.dz_1 * g.dz_1;
  return g;
}

#define F(g, f, ix, iy, iz, k) \
  (f)[(long)(k) * (g)->mw + ((long)(iz) * (g)->my + (iy)) * (g)->mx + (ix)]

/* ------------------------------------------------------------------ */
/* bc_sym_z (boundcond.f90:3842-3900) for one variable a (full x,y extent),
 * bottom then top, with the codes of boundconds_z ('s' :1486-1488, 'a'
 * :1498-1500, 'a2' :1501-1503). */
static void bc_sym_z(double* a, int mx, int my, int nz, int code, int top) {
  const long pl = (long)mx * my;
  const int nb = top ? NG + nz - 1 : NG; /* n2 or n1 */
  for (int i = 1; i <= NG; ++i) {
    double* gh = a + (long)(top ? nb + i : nb - i) * pl;
    const double* in = a + (long)(top ? nb - i : nb + i) * pl;
    const double* bd = a + (long)nb * pl;
    for (long q = 0; q < pl; ++q) {
      if (code == PCO_BC_S) gh[q] = in[q];
      else if (code == PCO_BC_A) gh[q] = -in[q];
      else if (top) gh[q] = bd[q] + (bd[q] - in[q]);  /* f(n2)+(f(n2)+sgn*f(n2-i)) */
      else gh[q] = 2 * bd[q] - in[q];                  /* 2*f(n1)+sgn*f(n1+i) */
    }
  }
  if (code == PCO_BC_A) memset(a + (long)nb * pl, 0, sizeof(double) * pl); /* set bdry value=0 */
}

/* 'c1' and 'cT' on ss (log density, no reference state), after the ghost
 * zones of lnrho are set (lnrho precedes ss in the variable loop of
 * boundconds_z).  Plane nb is n1 (bottom) or n2 (top); dz2[i] is
 * dz2_bound(-i) = 2*(z(n1+i)-z(n1)) or dz2_bound(i) = 2*(z(n2)-z(n2-i))
 * (grid.f90:3149-3151).
 * 'c1' (bc_ss_flux, entropy_bcs.f90:93-235, the K-profile branch):
 *   cs2 = cs20*exp(gamma_m1*(lnrho(nb)-lnrho0)+cv1*ss(nb)), tmp = cp*FbotKbot/cs2,
 *   bottom ss(n1-i) = ss(n1+i) + (cp-cv)*(dlnrho + dz2_bound(-i)*tmp),
 *          dlnrho = lnrho(n1+i) - lnrho(n1-i) (getdlnrho_z, density_methods.f90:409-426);
 *   top    ss(n2+i) = ss(n2-i) + (cp-cv)*(-dlnrho - dz2_bound(i)*tmp),
 *          dlnrho = lnrho(n2+i) - lnrho(n2-i), tmp with FtopKtop.
 * 'cT' (bc_ss_temp_z, entropy_bcs.f90:1329-1457): tmp = 2*cv*log(cs2b/cs20),
 *   ss(nb) = 0.5*tmp - (cp-cv)*(lnrho(nb)-lnrho0),
 *   ss(nb-+i) = -ss(nb+-i) + tmp - (cp-cv)*(lnrho(nb+-i)+lnrho(nb-+i)-2*lnrho0). */
static void bc_ss_z(double* ss, const double* lr, const pco_params* p, int code, int top) {
  const int mx = p->nx + 2 * NG, my = p->ny + 2 * NG;
  const long pl = (long)mx * my;
  const int nb = top ? NG + p->nz - 1 : NG;
  double e[8];
  pco_eos_constants(p, e);
  const double cs20 = e[0], gamma_m1 = e[1], cv = e[4], cv1 = e[5], lnrho0 = e[6], cp = p->cp;
  const double dz = p->Lz / (p->nz - 1);
  double zc[2 * NG + 1]; /* z(nb + j), j = -NG..NG */
  for (int j = -NG; j <= NG; ++j) zc[j + NG] = p->z0 + (nb - NG + j) * dz;
  double* sb = ss + (long)nb * pl;
  const double* lb = lr + (long)nb * pl;
  if (code == PCO_BC_C1) {
    const double hk = top ? p->hcond0 * p->hcond2 : p->hcond0 * p->hcond1;
    const double fk = hk != 0.0 ? (top ? p->ftop : p->fbot) / hk : 0.0; /* FbotKbot / FtopKtop (:851-873) */
    for (long q = 0; q < pl; ++q) {
      const double cs2 = cs20 * exp(gamma_m1 * (lb[q] - lnrho0) + cv1 * sb[q]);
      const double tmp = cp * fk / cs2;
      for (int i = 1; i <= NG; ++i) {
        const long in = q + (long)(top ? -i : i) * pl, gh = q + (long)(top ? i : -i) * pl;
        if (!top) {
          const double dz2 = 2. * (zc[NG + i] - zc[NG]);
          const double dlnrho = lb[in] - lb[gh];
          sb[gh] = sb[in] + (cp - cv) * (dlnrho + dz2 * tmp);
        } else {
          const double dz2 = 2. * (zc[NG] - zc[NG - i]);
          const double dlnrho = lb[gh] - lb[in];
          sb[gh] = sb[in] + (cp - cv) * (-dlnrho - dz2 * tmp);
        }
      }
    }
  } else if (code == PCO_BC_CT) {
    const double cs2b = top ? p->cs2top : p->cs2bot;
    const double tmp = 2 * cv * log(cs2b / cs20);
    for (long q = 0; q < pl; ++q) {
      sb[q] = 0.5 * tmp - (cp - cv) * (lb[q] - lnrho0);
      for (int i = 1; i <= NG; ++i) {
        const long in = q + (long)(top ? -i : i) * pl, gh = q + (long)(top ? i : -i) * pl;
        sb[gh] = -sb[in] + tmp - (cp - cv) * (lb[in] + lb[gh] - 2 * lnrho0);
      }
    }
  }
}

void pco_boundconds(double* f, const pco_params* p) {
  const int nvar = pco_nvar(p);
  pco_fill_ghosts(f, p->nx, p->ny, p->nz, nvar);
  if (p->lperi_z) return;
  const int mx = p->nx + 2 * NG, my = p->ny + 2 * NG;
  const long mw = (long)mx * my * (p->nz + 2 * NG);
  const int iss = pco_iss(p);
  for (int k = 0; k < nvar; ++k)
    for (int top = 0; top < 2; ++top) {
      const int code = top ? p->bcz_top[k] : p->bcz_bot[k];
      if (code == PCO_BC_C1 || code == PCO_BC_CT) {
        if (k == iss) bc_ss_z(f + k * mw, f + 3 * mw, p, code, top);
      } else {
        bc_sym_z(f + k * mw, mx, my, p->nz, code, top);
      }
    }
}

/* ------------------------------------------------------------------ */
/* periodic ghost zones: boundcond.f90:2387-2409 (x), :2411-2433 (y),
 * :2509- (z); boundconds_x, _y, _z order at equ.f90:209-224.          */
void pco_fill_ghosts(double* f, int nx, int ny, int nz, int nvar) {
  int mx = nx + 2 * NG, my = ny + 2 * NG, mz = nz + 2 * NG;
  long mw = (long)mx * my * mz;
  for (int k = 0; k < nvar; ++k) {
    double* a = f + k * mw;
    /* x: f(1:l1-1)=f(l2i:l2), f(l2+1:mx)=f(l1:l1i), all m, n */
#pragma omp parallel for schedule(static)
    for (int iz = 0; iz < mz; ++iz)
      for (int iy = 0; iy < my; ++iy) {
        double* row = a + ((long)iz * my + iy) * mx;
        for (int g = 0; g < NG; ++g) {
          row[g] = row[nx + g];
          row[NG + nx + g] = row[NG + g];
        }
      }
    /* y over full x extent */
#pragma omp parallel for schedule(static)
    for (int iz = 0; iz < mz; ++iz)
      for (int g = 0; g < NG; ++g) {
        memcpy(a + ((long)iz * my + g) * mx, a + ((long)iz * my + ny + g) * mx, sizeof(double) * mx);
        memcpy(a + ((long)iz * my + NG + ny + g) * mx, a + ((long)iz * my + NG + g) * mx,
               sizeof(double) * mx);
      }
    /* z over full x,y extents */
    for (int g = 0; g < NG; ++g) {
      memcpy(a + (long)g * my * mx, a + (long)(nz + g) * my * mx, sizeof(double) * mx * my);
      memcpy(a + (long)(NG + nz + g) * my * mx, a + (long)(NG + g) * my * mx, sizeof(double) * mx * my);
    }
  }
}

/* ------------------------------------------------------------------ */
/* der_main, deriv.f90:89-171 (Cartesian, equidistant).  j in {0,1,2}. */
static void der(const geom* g, const double* f, int m, int n, int k, int j, double* df) {
  const int nx = g->nx;
  if (j == 0) {
    const double fac = DER_A * g->dx_1;
    const double* r = &F(g, f, NG, m, n, k);
    for (int l = 0; l < nx; ++l)
      df[l] = fac * (45.0 * (r[l + 1] - r[l - 1]) - 9.0 * (r[l + 2] - r[l - 2]) + (r[l + 3] - r[l - 3]));
  } else {
    const double facs = DER_A * (j == 1 ? g->dy_1 : g->dz_1);
    const long s = (j == 1) ? g->mx : (long)g->mx * g->my;
    const double* r = &F(g, f, NG, m, n, k);
    for (int l = 0; l < nx; ++l) {
      const double* c = r + l;
      df[l] = facs * (45.0 * (c[s] - c[-s]) - 9.0 * (c[2 * s] - c[-2 * s]) + (c[3 * s] - c[-3 * s]));
    }
  }
}

/* der2_main, deriv.f90:472-561. */
static void der2(const geom* g, const double* f, int m, int n, int k, int j, double* df2) {
  const int nx = g->nx;
  long s;
  double fac;
  if (j == 0) { s = 1; fac = g->dx_1 * g->dx_1; }
  else if (j == 1) { s = g->mx; fac = g->dy_1 * g->dy_1; }
  else { s = (long)g->mx * g->my; fac = g->dz_1 * g->dz_1; }
  const double* r = &F(g, f, NG, m, n, k);
  for (int l = 0; l < nx; ++l) {
    const double* c = r + l;
    df2[l] = fac * (DER2_C0 * c[0] + DER2_C1 * (c[s] + c[-s]) + DER2_C2 * (c[2 * s] + c[-2 * s]) +
                    DER2_C3 * (c[3 * s] + c[-3 * s]));
  }
}

/* der6_main with UPWIND=.true. (deriv.f90:936-1001): fac=(1/60)*dx_1,
 * df = fac*(-20 f + 15 (f+1 + f-1) - 6 (f+2 + f-2) + (f+3 + f-3)). */
static void der6_upw(const geom* g, const double* f, int m, int n, int k, int j, double* df) {
  const int nx = g->nx;
  long s;
  double fac;
  if (j == 0) { s = 1; fac = (1.0 / 60) * g->dx_1; }
  else if (j == 1) { s = g->mx; fac = (1.0 / 60) * g->dy_1; }
  else { s = (long)g->mx * g->my; fac = (1. / 60.) * g->dz_1; }
  const double* r = &F(g, f, NG, m, n, k);
  for (int l = 0; l < nx; ++l) {
    const double* c = r + l;
    df[l] = fac * (-(20.0 * c[0]) + 15.0 * (c[s] + c[-s]) - 6.0 * (c[2 * s] + c[-2 * s]) + (c[3 * s] + c[-3 * s]));
  }
}

/* doupwind (sub.f90:7541-7574) through calc_del6_for_upwind (:7491-7539):
 * ugradf -= sum_j |u_j| * der6_upw_j(f_k), the sum over j in order. */
static void doupwind(const geom* g, const double* f, int m, int n, int k, const double* uu, double* ugradf,
                     double* tmp) {
  const int nx = g->nx;
  double* d = tmp;  /* 3*nx */
  for (int j = 0; j < 3; ++j) {
    der6_upw(g, f, m, n, k, j, d + (long)j * nx);
    for (int l = 0; l < nx; ++l) d[(long)j * nx + l] = fabs(uu[(long)j * nx + l]) * d[(long)j * nx + l];
  }
  for (int l = 0; l < nx; ++l) {
    const double up = d[l] + d[nx + l] + d[2 * nx + l];  /* sum(del6f,2) */
    ugradf[l] = ugradf[l] - up;
  }
}

/* derij_main bidiagonal branch, deriv.f90:1400-1452.  (i,j) unordered,
 * i != j.  The operand order inside each bracket follows the reference,
 * which differs between the xy/xz and the yz branches. */
static void derij(const geom* g, const double* f, int m, int n, int k, int i, int j, double* df) {
  const int nx = g->nx;
  const long sx = 1, sy = g->mx, sz = (long)g->mx * g->my;
  const double* r = &F(g, f, NG, m, n, k);
  const int s = i + j; /* 0-based: xy=1, xz=2, yz=3 */
  if (s == 1 || s == 2) {
    /* xy: fac=(1./720.)*dx_1*dy_1 ; xz: fac=(1./720.)*dz_1*dx_1 */
    const double fac = (s == 1) ? (1.0 / 720.0) * g->dx_1 * g->dy_1 : (1.0 / 720.0) * g->dz_1 * g->dx_1;
    const long t = (s == 1) ? sy : sz; /* second direction */
    for (int l = 0; l < nx; ++l) {
      const double* c = r + l;
#define BX(q) (c[(q) * sx + (q) * t] - c[-(q) * sx + (q) * t] + c[-(q) * sx - (q) * t] - c[(q) * sx - (q) * t])
      df[l] = fac * (270.0 * BX(1) - 27.0 * BX(2) + 2.0 * BX(3));
#undef BX
    }
  } else {
    /* yz: facs=(1./720.)*dy_1(m)*dz_1(n); f(m+1,n+1)-f(m+1,n-1)+f(m-1,n-1)-f(m-1,n+1) */
    const double facs = (1.0 / 720.0) * g->dy_1 * g->dz_1;
    for (int l = 0; l < nx; ++l) {
      const double* c = r + l;
#define BY(q) (c[(q) * sy + (q) * sz] - c[(q) * sy - (q) * sz] + c[-(q) * sy - (q) * sz] - c[-(q) * sy + (q) * sz])
      df[l] = facs * (270.0 * BY(1) - 27.0 * BY(2) + 2.0 * BY(3));
#undef BY
    }
  }
}

/* ------------------------------------------------------------------ */
/* pencil case (the subset this physics requests)                      */
enum {
  P_U = 0,                       /* uu(3) */
  P_U2 = P_U + 3, P_UIJ = P_U2 + 1, P_DIVU = P_UIJ + 9, P_SIJ = P_DIVU + 1, P_SIJ2 = P_SIJ + 9,
  P_UGU = P_SIJ2 + 1, P_DEL2U = P_UGU + 3, P_GDIVU = P_DEL2U + 3, P_ADVUU = P_GDIVU + 3,
  P_LNRHO = P_ADVUU + 1, P_RHO1 = P_LNRHO + 1, P_GLNRHO = P_RHO1 + 1, P_UGLNRHO = P_GLNRHO + 3,
  P_SGLNRHO = P_UGLNRHO + 1, P_DEL2LNRHO = P_SGLNRHO + 3,
  P_CS2 = P_DEL2LNRHO + 1, P_SS = P_CS2 + 1, P_GSS = P_SS + 1, P_DEL2SS = P_GSS + 3,
  P_LNTT = P_DEL2SS + 1, P_TT1 = P_LNTT + 1, P_GLNTT = P_TT1 + 1, P_DEL2LNTT = P_GLNTT + 3,
  P_UGSS = P_DEL2LNTT + 1, P_FPRES = P_UGSS + 1, P_ADVCS2 = P_FPRES + 3,
  P_FVISC = P_ADVCS2 + 1, P_VISCHEAT = P_FVISC + 3,
  P_AA = P_VISCHEAT + 1, P_AIJ = P_AA + 3, P_BB = P_AIJ + 9, P_B2 = P_BB + 3, P_UXB = P_B2 + 1,
  P_D2A = P_UXB + 3 /* d2A(j,k,i) 27 */, P_BIJ = P_D2A + 27, P_DEL2A = P_BIJ + 9,
  P_CURLB = P_DEL2A + 3, P_JJ = P_CURLB + 3, P_J2 = P_JJ + 3, P_JXB = P_J2 + 1, P_JXBR = P_JXB + 3,
  P_ADVVA2 = P_JXBR + 3,
  /* time step accumulators (equ.f90 module variables, per pencil) */
  P_MAXADVEC = P_ADVVA2 + 1, P_ADVEC2 = P_MAXADVEC + 1, P_ADVECCS2 = P_ADVEC2 + 1,
  P_MAXDIFFUS = P_ADVECCS2 + 1, P_DT1MAX = P_MAXDIFFUS + 1,
  P_TMP = P_DT1MAX + 1, P_TMP2 = P_TMP + 1, P_FIJJ = P_TMP2 + 1, P_FJJI = P_FIJJ + 9,
  P_UPW = P_FJJI + 9 /* scratch of the upwind correction (3) */,
  P_COUNT = P_UPW + 3
};

typedef struct {
  double* a; /* P_COUNT * nx */
  int nx;
} pencils;

#define PC(pc, i) ((pc)->a + (long)(i) * (pc)->nx)
#define UIJ(pc, i, j) PC(pc, P_UIJ + 3 * (i) + (j))
#define SIJ(pc, i, j) PC(pc, P_SIJ + 3 * (i) + (j))
#define AIJ(pc, i, j) PC(pc, P_AIJ + 3 * (i) + (j))
#define BIJ(pc, i, j) PC(pc, P_BIJ + 3 * (i) + (j))
#define D2A(pc, j, k, i) PC(pc, P_D2A + 9 * (i) + 3 * (j) + (k))

typedef struct {
  geom g;
  const pco_params* p;
  double cs20, gamma_m1, gamma1, cp1, cv, cv1, lnrho0, lnTT0, mu01;
  int iss, iax, nvar;
  int lupdate_courant_dt;
} ctx;

/* gij, sub.f90:1505-1544 (nder=1): g(:,i,j) = d f_{k+i} / dx_j */
static void gij(const ctx* c, const double* f, int m, int n, int k, double* g0) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) der(&c->g, f, m, n, k + i, j, g0 + (long)(3 * i + j) * c->g.nx);
}

/* dot_mn, sub.f90:599-626: c=a1*b1; c=c+a2*b2; c=c+a3*b3 */
static void dot_mn(int nx, const double* a0, long sa, const double* b0, long sb, double* c) {
  for (int l = 0; l < nx; ++l) {
    double s = a0[l] * b0[l];
    s = s + a0[sa + l] * b0[sb + l];
    s = s + a0[2 * sa + l] * b0[2 * sb + l];
    c[l] = s;
  }
}

/* dot2_mn, sub.f90:872-905: b=a(:,1)**2+a(:,2)**2+a(:,3)**2 */
static void dot2_mn(int nx, const double* a0, long sa, double* b) {
  for (int l = 0; l < nx; ++l) {
    double a1 = a0[l], a2 = a0[sa + l], a3 = a0[2 * sa + l];
    b[l] = a1 * a1 + a2 * a2 + a3 * a3;
  }
}

/* cross_mn, sub.f90:1456-1470 */
static void cross_mn(int nx, const double* a, const double* b, double* cc) {
  for (int l = 0; l < nx; ++l) {
    double a1 = a[l], a2 = a[nx + l], a3 = a[2 * nx + l];
    double b1 = b[l], b2 = b[nx + l], b3 = b[2 * nx + l];
    cc[l] = a2 * b3 - a3 * b2;
    cc[nx + l] = a3 * b1 - a1 * b3;
    cc[2 * nx + l] = a1 * b2 - a2 * b1;
  }
}

/* ------------------------------------------------------------------ */
/* calc_all_pencils (equ.f90:887-993) restricted to this physics.       */
static void calc_pencils(const ctx* c, const double* f, int m, int n, pencils* pc) {
  const geom* g = &c->g;
  const int nx = g->nx;
  const pco_params* p = c->p;
  /* ---- calc_pencils_hydro_nonlinear, hydro.f90:3925-4000 ---- */
  for (int i = 0; i < 3; ++i) memcpy(PC(pc, P_U + i), &F(g, f, NG, m, n, i), sizeof(double) * nx);
  dot2_mn(nx, PC(pc, P_U), nx, PC(pc, P_U2)); /* u2 */
  gij(c, f, m, n, 0, UIJ(pc, 0, 0));          /* uij: hydro.f90:3603-3607 */
  {
    /* del2v_etc(f,iuu,DEL2=del2u,GRADDIV=graddivu), sub.f90:2610-2749:
     * fijj(:,i,j)=der2(u_i, j); fjji(:,i,j)=derij(u_j, j, i), with the
     * i==j entry equal to der2(u_i,i) (derij leaves tmp untouched there). */
    double *del2u = PC(pc, P_DEL2U), *gdiv = PC(pc, P_GDIVU);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) der2(g, f, m, n, i, j, PC(pc, P_FIJJ + 3 * i + j));
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double* t = PC(pc, P_FJJI + 3 * i + j);
        if (i == j) memcpy(t, PC(pc, P_FIJJ + 3 * i + j), sizeof(double) * nx);
        else derij(g, f, m, n, j, j, i, t);
      }
    for (int i = 0; i < 3; ++i) {
      const double *a0 = PC(pc, P_FIJJ + 3 * i), *a1 = PC(pc, P_FIJJ + 3 * i + 1), *a2 = PC(pc, P_FIJJ + 3 * i + 2);
      const double *b0 = PC(pc, P_FJJI + 3 * i), *b1 = PC(pc, P_FJJI + 3 * i + 1), *b2 = PC(pc, P_FJJI + 3 * i + 2);
      for (int l = 0; l < nx; ++l) {
        del2u[i * nx + l] = a0[l] + a1[l] + a2[l];
        gdiv[i * nx + l] = b0[l] + b1[l] + b2[l];
      }
    }
  }
  /* ugu: u_dot_grad_vec (sub.f90:3482-3534) -> dot_mn(uu, uij(:,j,:)) */
  for (int j = 0; j < 3; ++j) dot_mn(nx, PC(pc, P_U), nx, UIJ(pc, j, 0), nx, PC(pc, P_UGU + j));
  /* lupw_uu: u_dot_grad_scl(...,UPWIND=lupw_uu) per component (sub.f90:3507, :3680) */
  if (p->lupw_uu)
    for (int j = 0; j < 3; ++j) doupwind(g, f, m, n, j, PC(pc, P_U), PC(pc, P_UGU + j), PC(pc, P_UPW));
  /* divu: div_mn, sub.f90:1981-2012 */
  for (int l = 0; l < nx; ++l) PC(pc, P_DIVU)[l] = UIJ(pc, 0, 0)[l] + UIJ(pc, 1, 1)[l] + UIJ(pc, 2, 2)[l];
  /* sij: traceless_strain, sub.f90:8922-8989 */
  for (int j = 0; j < 3; ++j) {
    for (int l = 0; l < nx; ++l) SIJ(pc, j, j)[l] = UIJ(pc, j, j)[l] - (1.0 / 3.0) * PC(pc, P_DIVU)[l];
    for (int i = j + 1; i < 3; ++i)
      for (int l = 0; l < nx; ++l) {
        double s = .5 * (UIJ(pc, i, j)[l] + UIJ(pc, j, i)[l]);
        SIJ(pc, i, j)[l] = s;
        SIJ(pc, j, i)[l] = s;
      }
  }
  /* sij2: multm2_sym_mn, sub.f90:1207-1226 */
  for (int l = 0; l < nx; ++l) {
    double b = SIJ(pc, 0, 0)[l] * SIJ(pc, 0, 0)[l];
    for (int i = 1; i < 3; ++i) {
      b = b + SIJ(pc, i, i)[l] * SIJ(pc, i, i)[l];
      for (int j = 0; j < i; ++j) b = b + 2 * (SIJ(pc, i, j)[l] * SIJ(pc, i, j)[l]);
    }
    PC(pc, P_SIJ2)[l] = b;
  }
  /* advec_uu = sum(abs(uu)*dline_1,2), hydro.f90:3500; maxadvec=max(maxadvec,advec_uu) :3530 */
  if (c->lupdate_courant_dt) {
    for (int l = 0; l < nx; ++l) {
      double a = fabs(PC(pc, P_U)[l]) * g->dx_1 + fabs(PC(pc, P_U + 1)[l]) * g->dy_1 +
                 fabs(PC(pc, P_U + 2)[l]) * g->dz_1;
      PC(pc, P_ADVUU)[l] = a;
      double* mx = PC(pc, P_MAXADVEC);
      mx[l] = mx[l] > a ? mx[l] : a;
    }
  }

  /* ---- calc_pencils_log_density_pnc, density.f90:2727-2811 ---- */
  memcpy(PC(pc, P_LNRHO), &F(g, f, NG, m, n, 3), sizeof(double) * nx);
  for (int l = 0; l < nx; ++l) PC(pc, P_RHO1)[l] = exp(-PC(pc, P_LNRHO)[l]);
  for (int j = 0; j < 3; ++j) der(g, f, m, n, 3, j, PC(pc, P_GLNRHO + j));
  dot_mn(nx, PC(pc, P_U), nx, PC(pc, P_GLNRHO), nx, PC(pc, P_UGLNRHO));
  /* lupw_lnrho: u_dot_grad(f,ilnrho,glnrho,uu,uglnrho,UPWIND=lupw_lnrho) (density.f90:2772-2773) */
  if (p->lupw_lnrho) doupwind(g, f, m, n, 3, PC(pc, P_U), PC(pc, P_UGLNRHO), PC(pc, P_UPW));
  if (p->lentropy) {
    /* del2(f,ilnrho): del2_main, sub.f90:2304-2336 */
    double* t = PC(pc, P_TMP);
    double* d = PC(pc, P_DEL2LNRHO);
    der2(g, f, m, n, 3, 0, d);
    der2(g, f, m, n, 3, 1, t);
    double* t2 = PC(pc, P_TMP2);
    der2(g, f, m, n, 3, 2, t2);
    for (int l = 0; l < nx; ++l) d[l] = d[l] + t[l] + t2[l];
  }
  /* sglnrho = multmv(sij,glnrho), density.f90:2805 / sub.f90:1228-1264 */
  for (int i = 0; i < 3; ++i)
    for (int l = 0; l < nx; ++l) {
      double t = SIJ(pc, i, 0)[l] * PC(pc, P_GLNRHO)[l];
      t = t + SIJ(pc, i, 1)[l] * PC(pc, P_GLNRHO + 1)[l];
      t = t + SIJ(pc, i, 2)[l] * PC(pc, P_GLNRHO + 2)[l];
      PC(pc, P_SGLNRHO + i)[l] = t;
    }

  /* ---- calc_pencils_eos_pencpar, eos_idealgas.f90:789-1161 ---- */
  if (!p->lentropy) {
    if (p->gamma == 1.0) {
      /* ilnrho_cs2, isothermal (:969-1000): cs2=cs20 */
      for (int l = 0; l < nx; ++l) PC(pc, P_CS2)[l] = c->cs20;
    } else {
      /* ilnrho_ss, isentropic/polytropic (:825-833): cs2=cs20*exp(gamma_m1*(lnrho-lnrho0)) */
      for (int l = 0; l < nx; ++l)
        PC(pc, P_CS2)[l] = c->cs20 * exp(c->gamma_m1 * (PC(pc, P_LNRHO)[l] - c->lnrho0));
    }
  } else {
    /* ilnrho_ss general branch (:851-891) */
    const int iss = c->iss;
    memcpy(PC(pc, P_SS), &F(g, f, NG, m, n, iss), sizeof(double) * nx);
    for (int j = 0; j < 3; ++j) der(g, f, m, n, iss, j, PC(pc, P_GSS + j));
    {
      double* d = PC(pc, P_DEL2SS);
      double* t = PC(pc, P_TMP);
      double* t2 = PC(pc, P_TMP2);
      der2(g, f, m, n, iss, 0, d);
      der2(g, f, m, n, iss, 1, t);
      der2(g, f, m, n, iss, 2, t2);
      for (int l = 0; l < nx; ++l) d[l] = d[l] + t[l] + t2[l];
    }
    for (int l = 0; l < nx; ++l) {
      double ss = PC(pc, P_SS)[l], lnrho = PC(pc, P_LNRHO)[l];
      PC(pc, P_CS2)[l] = c->cs20 * exp(c->cv1 * ss + c->gamma_m1 * (lnrho - c->lnrho0));
      double lnTT = c->lnTT0 + c->cv1 * ss + c->gamma_m1 * (lnrho - c->lnrho0);
      PC(pc, P_LNTT)[l] = lnTT;
      PC(pc, P_TT1)[l] = exp(-lnTT);
      for (int j = 0; j < 3; ++j)
        PC(pc, P_GLNTT + j)[l] = c->gamma_m1 * PC(pc, P_GLNRHO + j)[l] + c->cv1 * PC(pc, P_GSS + j)[l];
      PC(pc, P_DEL2LNTT)[l] = c->gamma_m1 * PC(pc, P_DEL2LNRHO)[l] + c->cv1 * PC(pc, P_DEL2SS)[l];
    }
  }

  /* ---- calc_pencils_energy ---- */
  if (!p->lentropy) {
    /* noentropy.f90:252-300: fpres=-cs2*glnrho; advec_cs2=cs2*dxyz_2 (:316) */
    for (int j = 0; j < 3; ++j)
      for (int l = 0; l < nx; ++l) PC(pc, P_FPRES + j)[l] = -PC(pc, P_CS2)[l] * PC(pc, P_GLNRHO + j)[l];
  } else {
    /* entropy.f90:3554-3670: ugss=u.gss; fpres=-cs2*(glnrho+glnTT)*gamma1 */
    dot_mn(nx, PC(pc, P_U), nx, PC(pc, P_GSS), nx, PC(pc, P_UGSS));
    /* u_dot_grad(f,iss,gss,uu,ugss,UPWIND=lupw_ss) (:3576) */
    if (p->lupw_ss) doupwind(g, f, m, n, c->iss, PC(pc, P_U), PC(pc, P_UGSS), PC(pc, P_UPW));
    for (int j = 0; j < 3; ++j)
      for (int l = 0; l < nx; ++l)
        PC(pc, P_FPRES + j)[l] =
            -(PC(pc, P_CS2)[l] * (PC(pc, P_GLNRHO + j)[l] + PC(pc, P_GLNTT + j)[l]) * c->gamma1);
  }
  if (c->lupdate_courant_dt)
    for (int l = 0; l < nx; ++l) PC(pc, P_ADVCS2)[l] = PC(pc, P_CS2)[l] * g->dxyz_2;

  /* ---- calc_pencils_viscosity, nu-const: viscosity.f90:1633-1663 ---- */
  for (int j = 0; j < 3; ++j)
    for (int l = 0; l < nx; ++l)
      PC(pc, P_FVISC + j)[l] = 0.0 + p->nu * (PC(pc, P_DEL2U + j)[l] + 2. * PC(pc, P_SGLNRHO + j)[l] +
                                              (1. / 3.) * PC(pc, P_GDIVU + j)[l]);
  if (p->lentropy)
    for (int l = 0; l < nx; ++l) PC(pc, P_VISCHEAT)[l] = 0.0 + 2 * p->nu * PC(pc, P_SIJ2)[l];
  if (c->lupdate_courant_dt) {
    /* diffus_nu = diffus_total*dxyz_2 (:2435-2436) */
    double* md = PC(pc, P_MAXDIFFUS);
    for (int l = 0; l < nx; ++l) {
      double d = p->nu * g->dxyz_2;
      md[l] = md[l] > d ? md[l] : d;
    }
  }

  /* ---- calc_pencils_magnetic_pencpar, magnetic.f90:4279-5225 ---- */
  if (p->lmagnetic) {
    const int iax = c->iax;
    for (int i = 0; i < 3; ++i) memcpy(PC(pc, P_AA + i), &F(g, f, NG, m, n, iax + i), sizeof(double) * nx);
    gij(c, f, m, n, iax, AIJ(pc, 0, 0)); /* aij */
    /* bb = curl_mn(aij) (sub.f90:2041-2097) */
    for (int l = 0; l < nx; ++l) {
      PC(pc, P_BB)[l] = AIJ(pc, 2, 1)[l] - AIJ(pc, 1, 2)[l];
      PC(pc, P_BB + 1)[l] = AIJ(pc, 0, 2)[l] - AIJ(pc, 2, 0)[l];
      PC(pc, P_BB + 2)[l] = AIJ(pc, 1, 0)[l] - AIJ(pc, 0, 1)[l];
    }
    dot2_mn(nx, PC(pc, P_BB), nx, PC(pc, P_B2));                /* b2 (:4419) */
    cross_mn(nx, PC(pc, P_U), PC(pc, P_BB), PC(pc, P_UXB));     /* uxb (:4455) */
    /* gij_etc(f,iaa,BIJ=bij,DEL2=del2a), sub.f90:2945-3113 */
    for (int i = 0; i < 3; ++i) {
      for (int j = 0; j < 3; ++j) der2(g, f, m, n, iax + i, j, D2A(pc, j, j, i));
      derij(g, f, m, n, iax + i, 1, 2, D2A(pc, 1, 2, i));
      memcpy(D2A(pc, 2, 1, i), D2A(pc, 1, 2, i), sizeof(double) * nx);
      derij(g, f, m, n, iax + i, 2, 0, D2A(pc, 2, 0, i));
      memcpy(D2A(pc, 0, 2, i), D2A(pc, 2, 0, i), sizeof(double) * nx);
      derij(g, f, m, n, iax + i, 0, 1, D2A(pc, 0, 1, i));
      memcpy(D2A(pc, 1, 0, i), D2A(pc, 0, 1, i), sizeof(double) * nx);
    }
    for (int j = 0; j < 3; ++j)
      for (int l = 0; l < nx; ++l) {
        BIJ(pc, 0, j)[l] = D2A(pc, 1, j, 2)[l] - D2A(pc, 2, j, 1)[l];
        BIJ(pc, 1, j)[l] = D2A(pc, 2, j, 0)[l] - D2A(pc, 0, j, 2)[l];
        BIJ(pc, 2, j)[l] = D2A(pc, 0, j, 1)[l] - D2A(pc, 1, j, 0)[l];
      }
    for (int i = 0; i < 3; ++i)
      for (int l = 0; l < nx; ++l)
        PC(pc, P_DEL2A + i)[l] = D2A(pc, 0, 0, i)[l] + D2A(pc, 1, 1, i)[l] + D2A(pc, 2, 2, i)[l];
    /* curlb = curl_mn(bij) (:4493); jj = mu01*curlb (:4735) */
    for (int l = 0; l < nx; ++l) {
      PC(pc, P_CURLB)[l] = BIJ(pc, 2, 1)[l] - BIJ(pc, 1, 2)[l];
      PC(pc, P_CURLB + 1)[l] = BIJ(pc, 0, 2)[l] - BIJ(pc, 2, 0)[l];
      PC(pc, P_CURLB + 2)[l] = BIJ(pc, 1, 0)[l] - BIJ(pc, 0, 1)[l];
    }
    for (int i = 0; i < 3; ++i)
      for (int l = 0; l < nx; ++l) PC(pc, P_JJ + i)[l] = c->mu01 * PC(pc, P_CURLB + i)[l];
    if (p->lentropy) dot2_mn(nx, PC(pc, P_JJ), nx, PC(pc, P_J2));
    /* jxb = cross_mn(jj,bb) (:4848); jxbr = rho1_jxb*jxb (:4888) */
    cross_mn(nx, PC(pc, P_JJ), PC(pc, P_BB), PC(pc, P_JXB));
    for (int i = 0; i < 3; ++i)
      for (int l = 0; l < nx; ++l) PC(pc, P_JXBR + i)[l] = PC(pc, P_RHO1)[l] * PC(pc, P_JXB + i)[l];
    if (c->lupdate_courant_dt) {
      /* advec_va2 = sum((bb*dline_1)**2,2)*mu01*rho1_jxb (:5180); advec2 += (:5217) */
      for (int l = 0; l < nx; ++l) {
        double b1 = PC(pc, P_BB)[l] * g->dx_1, b2 = PC(pc, P_BB + 1)[l] * g->dy_1,
               b3 = PC(pc, P_BB + 2)[l] * g->dz_1;
        double s = b1 * b1 + b2 * b2 + b3 * b3;
        PC(pc, P_ADVVA2)[l] = s * c->mu01 * PC(pc, P_RHO1)[l];
        PC(pc, P_ADVEC2)[l] = PC(pc, P_ADVEC2)[l] + PC(pc, P_ADVVA2)[l];
      }
    }
  }
}

/* ------------------------------------------------------------------ */
/* pi (cparam.f90), step and der_step (sub.f90:4975-5015), tini = 5*tiny
 * (cparam.f90:129). */
static const double PCO_PI = 3.14159265358979323846264338327950;
static const double TINI = 5 * 2.2250738585072014e-308;
static double step_fn(double x, double x0, double w) { return 0.5 * (1 + tanh((x - x0) / (w + TINI))); }
static double der_step_fn(double x, double x0, double w) {
  const double arg = fabs((x - x0) / (w + TINI));
  if (fabs(arg) >= 8.) return 2. / w * exp(-2. * fabs(arg));
  const double c = cosh(arg);
  return 0.5 / (w * (c * c));
}

/* z profiles at interior plane n: hcond and dlnhcond of get_gravz_heatcond
 * (entropy.f90:7741-7769), profz_heat (:1374-1386: luminosity times a
 * half-Gaussian at the bottom normalised by the horizontal area), and
 * profz_cool (:1319-1368: 'gaussian' about the top, or 'step' at z2, times
 * cool).  z(n) = xyz0(3) + n*dz with dz = Lz/(nz-1) (non-periodic z). */
void pco_layer_profile(const pco_params* p, int n, double out[4]) {
  const double dz = p->lperi_z ? p->Lz / p->nz : p->Lz / (p->nz - 1);
  const double z = p->z0 + (p->lperi_z ? n + 0.5 : (double)n) * dz;
  out[0] = out[1] = out[2] = out[3] = 0.0;
  if (p->lheatc_kprof) {
    double hc = 1. + (p->hcond1 - 1.) * step_fn(z, p->z1, -p->widthss) + (p->hcond2 - 1.) * step_fn(z, p->z2, p->widthss);
    const double dl = ((p->hcond1 - 1.) * der_step_fn(z, p->z1, -p->widthss) +
                       (p->hcond2 - 1.) * der_step_fn(z, p->z2, p->widthss)) / hc;
    out[0] = p->hcond0 * hc;
    out[1] = dl;
  }
  if (p->luminosity != 0.0) {
    const double t = (z - p->z0) / p->wheat;
    double h = p->luminosity * exp(-0.5 * (t * t)) / (sqrt(PCO_PI / 2.) * p->wheat);
    if (p->ny == 1 && p->nx > 1) h = h / p->Lx;
    else if (p->nx == 1 && p->ny > 1) h = h / p->Ly;
    else if (p->nx > 1 && p->ny > 1) h = h / (p->Lx * p->Ly);
    out[2] = h;
  }
  if (p->cool != 0.0) {
    double c;
    if (p->cooling_profile == 1) {
      c = step_fn(z, p->z2, p->wcool);
    } else {
      const double t = ((p->z0 + p->Lz) - z) / p->wcool;
      c = exp(-0.5 * (t * t));
    }
    out[3] = p->cool * c;
  }
}

/* ------------------------------------------------------------------ */
/* calc_all_rhs, equ.f90:1223-1366: duu_dt, dlnrho_dt, denergy_dt, daa_dt */
static void calc_rhs(const ctx* c, double* df, int m, int n, pencils* pc) {
  const geom* g = &c->g;
  const int nx = g->nx;
  const pco_params* p = c->p;
#define DF(k) (&F(g, df, NG, m, n, (k)))
  /* duu_dt: advec_uu -ugu (hydro.f90:4384), Coriolis force (:4494),
   * calc_viscous_force +fvisc (:4518, viscosity.f90:2823) */
  for (int j = 0; j < 3; ++j) {
    double* d = DF(j);
    for (int l = 0; l < nx; ++l) d[l] = d[l] - PC(pc, P_UGU + j)[l];
  }
  if (p->omega != 0.0) {
    /* coriolis_cartesian (hydro.f90:6339-6413): theta = 0: c2 = 2*Omega,
     * du_x += c2*u_y, du_y -= c2*u_x; else Omega = (-sin theta, 0, cos theta),
     * c2 = 2*Omega*cos(theta*dtor), s2 = -2*Omega*sin(theta*dtor). */
    double* dux = DF(0);
    double* duy = DF(1);
    double* duz = DF(2);
    const double* ux = PC(pc, P_U);
    const double* uy = PC(pc, P_U + 1);
    const double* uz = PC(pc, P_U + 2);
    if (p->theta == 0.0) {
      const double c2 = 2 * p->omega;
      for (int l = 0; l < nx; ++l) {
        dux[l] = dux[l] + c2 * uy[l];
        duy[l] = duy[l] - c2 * ux[l];
      }
    } else {
      const double dtor = PCO_PI / 180.;
      const double c2 = 2 * p->omega * cos(p->theta * dtor), s2 = -2 * p->omega * sin(p->theta * dtor);
      for (int l = 0; l < nx; ++l) {
        dux[l] = dux[l] + c2 * uy[l];
        duy[l] = duy[l] - c2 * ux[l] + s2 * uz[l];
        duz[l] = duz[l] - s2 * uy[l];
      }
    }
  }
  for (int j = 0; j < 3; ++j) {
    double* d = DF(j);
    for (int l = 0; l < nx; ++l) d[l] = d[l] + PC(pc, P_FVISC + j)[l];
  }
  /* dlnrho_dt -> continuity_eq: density_rhs = -divu - uglnrho (density.f90:2929-2930, :3059) */
  {
    double* d = DF(3);
    for (int l = 0; l < nx; ++l) {
      double density_rhs = -PC(pc, P_DIVU)[l];
      density_rhs = density_rhs - 1.0 * PC(pc, P_UGLNRHO)[l];
      d[l] = d[l] + density_rhs;
    }
  }
  /* denergy_dt */
  if (!p->lentropy) {
    /* noentropy.f90:381: df(uu) += fpres; advec_cs2=max(advec_cs2,p%advec_cs2) (:396) */
    for (int j = 0; j < 3; ++j) {
      double* d = DF(j);
      for (int l = 0; l < nx; ++l) d[l] = d[l] + PC(pc, P_FPRES + j)[l];
    }
    if (c->lupdate_courant_dt)
      for (int l = 0; l < nx; ++l) {
        double* a = PC(pc, P_ADVECCS2);
        a[l] = a[l] > PC(pc, P_ADVCS2)[l] ? a[l] : PC(pc, P_ADVCS2)[l];
      }
  } else {
    /* entropy.f90: advec_cs2 = p%advec_cs2 set in calc_pencils_energy (:3650) */
    if (c->lupdate_courant_dt) memcpy(PC(pc, P_ADVECCS2), PC(pc, P_ADVCS2), sizeof(double) * nx);
    for (int j = 0; j < 3; ++j) {
      double* d = DF(j);
      for (int l = 0; l < nx; ++l) d[l] = d[l] + PC(pc, P_FPRES + j)[l];
    }
    double* ds = DF(c->iss);
    /* -ugss (:3758), viscous heat TT1*visc_heat (viscosity.f90:2764) */
    for (int l = 0; l < nx; ++l) ds[l] = ds[l] - PC(pc, P_UGSS)[l];
    for (int l = 0; l < nx; ++l) ds[l] = ds[l] + PC(pc, P_TT1)[l] * PC(pc, P_VISCHEAT)[l];
    double zl[4];
    pco_layer_profile(p, n - NG, zl);
    /* calc_heatcond (entropy.f90:3800, :5913-6151) with iheatcond='K-profile':
     * glnThcond = glnTT + glhc, glhc = (0, 0, dlnhcond(z));
     * thdiff = rho1*hcond*(del2lnTT + glnTT.glnThcond); chix = rho1*hcond*cv1 */
    if (p->lheatc_kprof)
      for (int l = 0; l < nx; ++l) {
        const double g1 = PC(pc, P_GLNTT)[l], g2v = PC(pc, P_GLNTT + 1)[l], g3 = PC(pc, P_GLNTT + 2)[l];
        double g2 = g1 * (g1 + 0.0);
        g2 = g2 + g2v * (g2v + 0.0);
        g2 = g2 + g3 * (g3 + zl[1]);
        const double thdiff = PC(pc, P_RHO1)[l] * zl[0] * (PC(pc, P_DEL2LNTT)[l] + g2);
        ds[l] = ds[l] + thdiff;
      }
    /* calc_heatcond_constchi (entropy.f90:4852-4947):
     * g2 = (glnrho+glnTT).glnTT ; thdiff = chi*cp*(del2lnTT+g2) */
    for (int l = 0; l < nx; ++l) {
      double a1 = PC(pc, P_GLNRHO)[l] + PC(pc, P_GLNTT)[l];
      double a2 = PC(pc, P_GLNRHO + 1)[l] + PC(pc, P_GLNTT + 1)[l];
      double a3 = PC(pc, P_GLNRHO + 2)[l] + PC(pc, P_GLNTT + 2)[l];
      double g2 = a1 * PC(pc, P_GLNTT)[l];
      g2 = g2 + a2 * PC(pc, P_GLNTT + 1)[l];
      g2 = g2 + a3 * PC(pc, P_GLNTT + 2)[l];
      double thdiff = p->chi * p->cp * (PC(pc, P_DEL2LNTT)[l] + g2);
      ds[l] = ds[l] + thdiff;
    }
    /* calc_heat_cool (:3839, :6387-6560) through get_heat_cool_gravz
     * (:6761-6828) when luminosity or cool is set: heat = profz_heat -
     * profz_cool*(cs2-cs2cool)/cs2cool (lcooling_to_cs2cool); df(ss) += TT1*rho1*heat */
    if (p->luminosity != 0.0 || p->cool != 0.0)
      for (int l = 0; l < nx; ++l) {
        double heat = 0.0;
        if (p->luminosity != 0.0) heat = zl[2];
        heat = heat - zl[3] * (PC(pc, P_CS2)[l] - p->cs2cool) / p->cs2cool;
        ds[l] = ds[l] + PC(pc, P_TT1)[l] * PC(pc, P_RHO1)[l] * heat;
      }
    if (c->lupdate_courant_dt) {
      double* md = PC(pc, P_MAXDIFFUS);
      for (int l = 0; l < nx; ++l) {
        double d = 0.0; /* diffus_chi */
        if (p->lheatc_kprof) d = d + PC(pc, P_RHO1)[l] * zl[0] * c->cv1 * g->dxyz_2; /* chix*dxyz_2 (:6094) */
        d = d + (p->gamma * p->chi) * g->dxyz_2;                                      /* (:4907) */
        md[l] = md[l] > d ? md[l] : d;
      }
    }
  }
  /* daa_dt (magnetic.f90:5392-6530) */
  if (p->lmagnetic) {
    /* Lorentz force df(uu) += jxbr (:5531) */
    for (int j = 0; j < 3; ++j) {
      double* d = DF(j);
      for (int l = 0; l < nx; ++l) d[l] = d[l] + PC(pc, P_JXBR + j)[l];
    }
    /* fres = eta*del2a (:5589); dAdt = 0 + uxb + fres (:6247); df(aa) += dAdt (:6524) */
    for (int j = 0; j < 3; ++j) {
      double* d = DF(c->iax + j);
      for (int l = 0; l < nx; ++l) {
        double fres = 0.0 + p->eta * PC(pc, P_DEL2A + j)[l];
        double dAdt = 0.0 + PC(pc, P_UXB + j)[l] + fres;
        d[l] = d[l] + dAdt;
      }
    }
    /* ohmic heating (:6143): df(ss) += eta_heat*mu0*j2*rho1*TT1 */
    if (p->lentropy) {
      double* ds = DF(c->iss);
      for (int l = 0; l < nx; ++l)
        ds[l] = ds[l] + p->eta * p->mu0 * PC(pc, P_J2)[l] * PC(pc, P_RHO1)[l] * PC(pc, P_TT1)[l];
    }
    /* diffus_eta = eta_total*dxyz_2 (:6541) */
    if (c->lupdate_courant_dt) {
      double* md = PC(pc, P_MAXDIFFUS);
      for (int l = 0; l < nx; ++l) {
        double d = p->eta * g->dxyz_2;
        md[l] = md[l] > d ? md[l] : d;
      }
    }
  }
  /* addgravity (gravity_simple.f90:1083-1162, called after daa_dt at equ.f90:1316):
   * df(iuz) += gg(:,3), gg(:,3) = gravz for gravz_profile='const' */
  if (p->gravz != 0.0) {
    double* d = DF(2);
    for (int l = 0; l < nx; ++l) d[l] = d[l] + p->gravz;
  }
#undef DF
}

/* calc_maxadvec + set_dt1_max, equ.f90:1858-1912 */
static void set_dt1_max(const ctx* c, pencils* pc) {
  const int nx = c->g.nx;
  const pco_params* p = c->p;
  for (int l = 0; l < nx; ++l) {
    double advec2 = PC(pc, P_ADVEC2)[l] + PC(pc, P_ADVECCS2)[l];
    double maxadvec = PC(pc, P_MAXADVEC)[l] + sqrt(advec2);
    maxadvec = maxadvec + sqrt(0.0); /* advec2_hypermesh */
    double dt1_advec = maxadvec / p->cdt;
    double dt1_diffus = PC(pc, P_MAXDIFFUS)[l] / p->cdtv + 0.0 / 0.03 + 0.0 / 0.01;
    double dt1_src = 0.0 / 1.0;
    double dt1_max_loc = sqrt(dt1_advec * dt1_advec + dt1_diffus * dt1_diffus + dt1_src * dt1_src);
    double* dm = PC(pc, P_DT1MAX);
    dm[l] = dm[l] > dt1_max_loc ? dm[l] : dt1_max_loc;
  }
}

static ctx make_ctx(const pco_params* p, int lupdate_courant_dt) {
  ctx c;
  c.g = make_geom(p);
  c.p = p;
  double e[8];
  pco_eos_constants(p, e);
  c.cs20 = e[0]; c.gamma_m1 = e[1]; c.gamma1 = e[2]; c.cp1 = e[3];
  c.cv = e[4]; c.cv1 = e[5]; c.lnrho0 = e[6]; c.lnTT0 = e[7];
  c.mu01 = 1. / p->mu0; /* magnetic.f90:1553 */
  c.iss = pco_iss(p); c.iax = pco_iax(p); c.nvar = pco_nvar(p);
  c.lupdate_courant_dt = lupdate_courant_dt;
  return c;
}

/* pde(), equ.f90:39-444 on the CPU path (rhs_cpu pencil loop :1443-1561). */
void pco_pde(double* f, double* df, const pco_params* p, int want_dt, double* dt1, double* diag,
             int nthreads) {
  ctx c = make_ctx(p, want_dt);
  const geom* g = &c.g;
  pco_boundconds(f, p);
  double dt1_all = 0.0, su2 = 0.0, mu2 = 0.0, sb2 = 0.0;
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
  nthreads = 1;
#endif
  const int nyz = g->ny * g->nz;
#pragma omp parallel num_threads(nthreads) reduction(max : dt1_all, mu2) reduction(+ : su2, sb2)
  {
    pencils pc;
    pc.nx = g->nx;
    pc.a = calloc((size_t)P_COUNT * g->nx, sizeof(double));
#pragma omp for schedule(static)
    for (int imn = 0; imn < nyz; ++imn) {
      const int m = NG + imn % g->ny, n = NG + imn / g->ny;
      if (c.lupdate_courant_dt) {
        memset(PC(&pc, P_MAXADVEC), 0, sizeof(double) * g->nx);
        memset(PC(&pc, P_ADVEC2), 0, sizeof(double) * g->nx);
        memset(PC(&pc, P_ADVECCS2), 0, sizeof(double) * g->nx);
        memset(PC(&pc, P_MAXDIFFUS), 0, sizeof(double) * g->nx);
        memset(PC(&pc, P_DT1MAX), 0, sizeof(double) * g->nx);
      }
      calc_pencils(&c, f, m, n, &pc);
      calc_rhs(&c, df, m, n, &pc);
      if (c.lupdate_courant_dt) {
        set_dt1_max(&c, &pc);
        for (int l = 0; l < g->nx; ++l) dt1_all = dt1_all > PC(&pc, P_DT1MAX)[l] ? dt1_all : PC(&pc, P_DT1MAX)[l];
      }
      if (diag) {
        /* sum_mn_name(p%u2,idiag_urms,lsqrt) hydro.f90:4709; max_mn_name umax :4743;
         * brms from b2 magnetic.f90:6879 */
        double s = 0.0, sm = 0.0;
        for (int l = 0; l < g->nx; ++l) {
          double u2 = PC(&pc, P_U2)[l];
          s += u2;
          mu2 = mu2 > u2 ? mu2 : u2;
          if (p->lmagnetic) sm += PC(&pc, P_B2)[l];
        }
        su2 += s;
        sb2 += sm;
      }
    }
    free(pc.a);
  }
  if (dt1) *dt1 = dt1_all;
  if (diag) {
    /* dVol_rel1 = 1/nwgrid on Cartesian equidistant grids (diagnostics.f90:274) */
    double nw = (double)g->nx * g->ny * g->nz;
    diag[0] = sqrt(su2 / nw);
    diag[1] = sqrt(mu2);
    diag[2] = p->lmagnetic ? sqrt(sb2 / nw) : 0.0;
  }
}

/* time_step / advance_substeps, timestep.f90:77-273 (CPU branch). */
void pco_advance(double* f, double* df, const pco_params* p, int nsteps, double* t, double* dt,
                 double* diag_series, int nthreads) {
  const geom g = make_geom(p);
  const int nvar = pco_nvar(p);
  const int ldt = (p->dt == 0.0);
  double dt_ = ldt ? *dt : p->dt;
#ifdef _OPENMP
  const int nth = nthreads > 0 ? nthreads : omp_get_max_threads();
#endif
  for (int it = 0; it < nsteps; ++it) {
    double dt_beta_ts[3];
    if (!ldt)
      for (int s = 0; s < 3; ++s) dt_beta_ts[s] = dt_ * BETA_TS[s];
    double ds = 0.0;
    for (int itsub = 0; itsub < 3; ++itsub) {
      const int lfirst = (itsub == 0);
      const long ntot = g.mw * nvar;
      if (lfirst) {
        memset(df, 0, sizeof(double) * ntot);
        ds = 0.0;
      } else {
        const double a = ALPHA_TS[itsub];
#pragma omp parallel for schedule(static) num_threads(nth)
        for (long q = 0; q < ntot; ++q) df[q] = a * df[q];
        ds = a * ds;
      }
      double dt1 = 0.0;
      pco_pde(f, df, p, lfirst && ldt, &dt1, (lfirst && diag_series) ? diag_series + 3L * it : NULL,
              nthreads);
      if (lfirst && ldt) dt_ = 1.0 / dt1; /* set_dt, sub.f90:4223-4254 */
      ds = ds + 1.0;
      if (ldt)
        for (int s = 0; s < 3; ++s) dt_beta_ts[s] = dt_ * BETA_TS[s];
      const double dtsub = ds * dt_beta_ts[itsub];
      const double db = dt_beta_ts[itsub];
#pragma omp parallel for collapse(2) schedule(static) num_threads(nth)
      for (int k = 0; k < nvar; ++k)
        for (int n = NG; n < NG + g.nz; ++n)
          for (int m = NG; m < NG + g.ny; ++m) {
            double* fr = &F(&g, f, NG, m, n, k);
            const double* dr = &F(&g, df, NG, m, n, k);
            for (int l = 0; l < g.nx; ++l) fr[l] = fr[l] + db * dr[l];
          }
      *t = *t + dtsub;
    }
  }
  *dt = dt_;
}

/* ------------------------------------------------------------------ */
/* single-stencil entry points for the stencil tests                   */
static void for_all_pencils(const double* f, const pco_params* p, double* out,
                            void (*fn)(const geom*, const double*, int, int, int, int, int, double*),
                            int k, int a, int b) {
  geom g = make_geom(p);
  for (int n = 0; n < g.nz; ++n)
    for (int m = 0; m < g.ny; ++m) fn(&g, f, m + NG, n + NG, k, a, b, out + ((long)n * g.ny + m) * g.nx);
}
static void der_w(const geom* g, const double* f, int m, int n, int k, int a, int b, double* o) {
  (void)b;
  der(g, f, m, n, k, a, o);
}
static void der2_w(const geom* g, const double* f, int m, int n, int k, int a, int b, double* o) {
  (void)b;
  der2(g, f, m, n, k, a, o);
}
void pco_der(const double* f, const pco_params* p, int k, int j, double* out) {
  for_all_pencils(f, p, out, der_w, k, j, 0);
}
void pco_der2(const double* f, const pco_params* p, int k, int j, double* out) {
  for_all_pencils(f, p, out, der2_w, k, j, 0);
}
void pco_derij(const double* f, const pco_params* p, int k, int i, int j, double* out) {
  for_all_pencils(f, p, out, derij, k, i, j);
}

/* ------------------------------------------------------------------ */
/* addforce, helical default path: see pc_oracle.h.                     */
void pco_addforce(double* f, const pco_params* p, const pco_forcing* fo) {
  const geom gg = make_geom(p);
  const geom* g = &gg;
  const int nx = g->nx, ny = g->ny, nz = g->nz;
  const double dx = p->Lx / p->nx, dy = p->Ly / p->ny;
  const double dz = p->lperi_z ? p->Lz / p->nz : p->Lz / (p->nz - 1);
  const double zsh = p->lperi_z ? 0.5 : 0.0; /* cell centres, or grid points from the boundary */
  double* fxr = malloc(sizeof(double) * 2 * (nx + ny + nz));
  double *fxi = fxr + nx, *fyr = fxi + nx, *fyi = fyr + ny, *fzr = fyi + ny, *fzi = fzr + nz;
  /* fxyz_coefs_hel: fx=exp(cmplx(0.,kk(1)*k1_ff*x+phase))*fact, fy, fz */
  for (int l = 0; l < nx; ++l) {
    const double x = fo->x0 + (l + 0.5) * dx;
    const double th = fo->kk[0] * fo->k1_ff * x + fo->phase;
    fxr[l] = cos(th) * fo->fact;
    fxi[l] = sin(th) * fo->fact;
  }
  for (int m = 0; m < ny; ++m) {
    const double y = fo->y0 + (m + 0.5) * dy;
    const double th = fo->kk[1] * fo->k1_ff * y;
    fyr[m] = cos(th);
    fyi[m] = sin(th);
  }
  for (int n = 0; n < nz; ++n) {
    const double z = fo->z0 + (n + zsh) * dz;
    const double th = fo->kk[2] * fo->k1_ff * z;
    fzr[n] = cos(th);
    fzi[n] = sin(th);
  }
#pragma omp parallel for schedule(static)
  for (int n = 0; n < nz; ++n)
    for (int m = 0; m < ny; ++m)
      for (int l = 0; l < nx; ++l) {
        /* fxyz = fx(l)*fy(m)*fz(n), complex products left to right */
        const double ar = fxr[l] * fyr[m] - fxi[l] * fyi[m];
        const double ai = fxr[l] * fyi[m] + fxi[l] * fyr[m];
        const double br = ar * fzr[n] - ai * fzi[n];
        const double bi = ar * fzi[n] + ai * fzr[n];
        for (int j = 0; j < 3; ++j) {
          /* forcing_rhs = force_ampl*fda(j)*cos(omega_ff*t)*real(cmplx(coef1(j),coef2(j))*fxyz) */
          const double re = fo->coef1[j] * br - fo->coef2[j] * bi;
          const double rhs = fo->fda[j] * re;
          F(g, f, NG + l, NG + m, NG + n, fo->target + j) += rhs;
        }
      }
  free(fxr);
}
