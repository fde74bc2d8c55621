// pcmx_physics.h -- pointwise RHS algebra of the fused kernels.
//
// Given the value, first, second and (bidiagonal) mixed derivatives of every
// evolved variable at one cell, adds the right-hand side of
//   src/hydro.f90 (duu_dt: -u.grad u), src/viscosity.f90 (nu-const force,
//   viscous heat), src/density.f90 (continuity for ln rho),
//   src/noentropy.f90 / src/entropy.f90 (pressure force, -u.grad s,
//   chi-const conduction) with src/eos_idealgas.f90, and src/magnetic.f90
//   (eta-const induction, J x B / rho, ohmic heat)
// to the 2N register w, in the order calc_all_rhs adds them
// (src/equ.f90:1268-1274: duu_dt, dlnrho_dt, denergy_dt, daa_dt), and
// evaluates the inverse Courant step of set_dt1_max (src/equ.f90:1858-1912).
// The operation grouping follows the reference expressions cited below so
// that the FP64 form agrees with the reference to rounding.
#pragma once

#include <hip/hip_runtime.h>

namespace pcmx {

// Constants of one box (host-computed once; identical for all kernels).
struct Phys {
  double dx_1, dy_1, dz_1, dxyz_2;
  // stencil prefactors, computed in the reference's operation order:
  double fac1[3];   // (1/60)*d?_1            deriv.f90:103,124,135,153
  double fac2[3];   // d?_1**2                deriv.f90:498,514,539
  double facxy;     // (1./720.)*dx_1*dy_1    deriv.f90:1409
  double facxz;     // (1./720.)*dz_1*dx_1    deriv.f90:1439
  double facyz;     // (1./720.)*dy_1*dz_1    deriv.f90:1424
  double nu, eta, chi, cp, mu0, mu01;
  double cs20, gamma, gamma_m1, gamma1, cv1, lnrho0, lnTT0;
  double cdt, cdtv;
  double gravz;     // constant vertical gravity (gravity_simple, gravz_profile='const'), or 0
  int polytropic;   // noentropy with gamma /= 1 (ilnrho_ss, leos_isentropic)
  // Layer physics of samples/conv-slab (kernels instantiated with LAY):
  // Coriolis coefficients c2 = 2 Omega cos(theta), s2 = -2 Omega sin(theta)
  // (coriolis_cartesian, hydro.f90:6339-6413); per z plane n the row
  // zlay[4n..4n+3] = {hcond, dlnhcond/dz, profz_heat, profz_cool}
  // (get_gravz_heatcond, entropy.f90:7741-7769; :1319-1386); cs2cool; kprof
  // for iheatcond='K-profile', heatcool when luminosity or cool is set.
  double cor_c2, cor_s2, cs2cool;
  const double* zlay;
  int kprof, heatcool;
};

// der2_coef*, deriv.f90:36-40 ('standard').
#define PCMX_D2C0 (-490.0 / 180.0)
#define PCMX_D2C1 (270.0 / 180.0)
#define PCMX_D2C2 (-27.0 / 180.0)
#define PCMX_D2C3 (2.0 / 180.0)

// Derivatives at one cell.  The second derivatives of uu, lnrho and ss enter
// the RHS only through del2 and graddiv (del2v_etc, sub.f90:2681-2701), so
// they are kept as those sums; a producer adds the terms in the reference's
// order: del2_i = (D_xx + D_yy) + D_zz of component i, and graddiv_i =
// sum over j = x, y, z of D_j D_i u_j, the u_x term first (so producers visit
// u_x, u_y, u_z in that order; the first term initialises the sum).
// mxa[pair][c]: pair 0 = xy, 1 = xz, 2 = yz, component c of A; only the
// entries the curl of B uses are filled.
template <typename T>
struct Derivs {
  T u[3], lnrho, ss, a[3];
  T uij[3][3];        // uij[i][j] = D_j u_i           (gij, sub.f90:1505)
  T glnrho[3], gss[3];
  T aij[3][3];        // D_j A_i
  T del2u[3], gdivu[3];
  T del2lnrho, del2ss;
  T d2a[3][3];        // D2_j A_i                      (gij_etc d2A(j,j,i))
  T mxa[3][3];
  // upwind corrections sum_j |u_j| D6_j f of u_x, u_y, u_z, lnrho, ss
  // (doupwind, sub.f90:7491-7574; D6_j = der6_main with UPWIND,
  // deriv.f90:936-1001), filled only with upwinding
  T upw[5];
};

// Upwinding masks (compile time): UPW_UU for lupw_uu (hydro.f90:3662),
// UPW_LNRHO for lupw_lnrho (density.f90:2772-2773), UPW_SS for lupw_ss
// (entropy.f90:3576).
constexpr int UPW_UU = 1, UPW_LNRHO = 2, UPW_SS = 4;

// der6_main with UPWIND=.true. along one direction from the symmetric sums
// s_k = f(+k) + f(-k): fac*(-20 f0 + 15 s1 - 6 s2 + s3) with fac = dx_1/60,
// the operand order of deriv.f90:990-993.
template <typename T>
__device__ __forceinline__ T der6_upwind(T fac, T c0, T s1, T s2, T s3) {
  return fac * (-((T)20.0 * c0) + (T)15.0 * s1 - (T)6.0 * s2 + s3);
}

// Adds the second derivatives of component i of uu to del2u and gdivu.
// mixed[] holds D_xy, D_xz, D_yz of u_i (entries not used by graddiv are
// ignored).  Must be called for i = 0, 1, 2 in that order.
template <typename T>
__device__ __forceinline__ void add_u_second(Derivs<T>& d, int i, T dxx, T dyy, T dzz, T mxy, T mxz, T myz) {
  d.del2u[i] = dxx + dyy + dzz;
  if (i == 0) {
    d.gdivu[0] = dxx;
    d.gdivu[1] = mxy;
    d.gdivu[2] = mxz;
  } else if (i == 1) {
    d.gdivu[0] = d.gdivu[0] + mxy;
    d.gdivu[1] = d.gdivu[1] + dyy;
    d.gdivu[2] = d.gdivu[2] + myz;
  } else {
    d.gdivu[0] = d.gdivu[0] + mxz;
    d.gdivu[1] = d.gdivu[1] + myz;
    d.gdivu[2] = d.gdivu[2] + dzz;
  }
}

template <typename T>
__device__ __forceinline__ T pexp(T x);
template <>
__device__ __forceinline__ double pexp<double>(double x) { return exp(x); }
template <>
__device__ __forceinline__ float pexp<float>(float x) { return __expf(x); }

template <typename T>
__device__ __forceinline__ T psqrt(T x);
template <>
__device__ __forceinline__ double psqrt<double>(double x) { return sqrt(x); }
template <>
__device__ __forceinline__ float psqrt<float>(float x) { return sqrtf(x); }

// multm2_sym_mn, sub.f90:1207-1226: b = s11^2; then per row i>1:
// b += s_ii^2; b += 2*s_ij^2 (j<i).
template <typename T>
__device__ __forceinline__ T d_sij2(const T (&s)[3][3]) {
  T b = s[0][0] * s[0][0];
  b = b + s[1][1] * s[1][1];
  b = b + (T)2 * (s[1][0] * s[1][0]);
  b = b + s[2][2] * s[2][2];
  b = b + (T)2 * (s[2][0] * s[2][0]);
  b = b + (T)2 * (s[2][1] * s[2][1]);
  return b;
}

// ---------------------------------------------------------------------------
// The pencils of the Courant condition (set_dt1_max, equ.f90:1858-1912),
// shared by the RHS kernels and the dt pass of Courant steps so that both
// evaluate them with the same arithmetic.

// Sound speed squared (eos_idealgas.f90:871, 969-1000).
template <typename T, bool ENT, bool POLY>
__device__ __forceinline__ T eos_cs2(T lnrho, T ss, const Phys& P) {
  if constexpr (ENT) return (T)P.cs20 * pexp<T>((T)P.cv1 * ss + (T)P.gamma_m1 * (lnrho - (T)P.lnrho0));
  else if constexpr (POLY) return (T)P.cs20 * pexp<T>((T)P.gamma_m1 * (lnrho - (T)P.lnrho0));
  else return (T)P.cs20;
}

// advec_uu (hydro.f90:3500).
template <typename T>
__device__ __forceinline__ T advec_uu_of(T u0, T u1, T u2, const Phys& P) {
  return fabs(u0) * (T)P.dx_1 + fabs(u1) * (T)P.dy_1 + fabs(u2) * (T)P.dz_1;
}

// advec_va2 (magnetic.f90:5180).
template <typename T>
__device__ __forceinline__ T advec_va2_of(const T (&bb)[3], T rho1, const Phys& P) {
  const T b1x = bb[0] * (T)P.dx_1, b1y = bb[1] * (T)P.dy_1, b1z = bb[2] * (T)P.dz_1;
  return (b1x * b1x + b1y * b1y + b1z * b1z) * (T)P.mu01 * rho1;
}

// dt1_max of one cell from the advective pencils (set_dt1_max).
template <typename T, bool MAG, bool ENT, bool LAY = false>
__device__ __forceinline__ double courant_dt1(T advec_uu, T cs2, T advec_va2, const Phys& P, T dchi = 0) {
  const T advec_cs2 = cs2 * (T)P.dxyz_2;
  const T advec2 = (MAG ? advec_va2 : (T)0) + advec_cs2;
  const T maxadvec = advec_uu + psqrt<T>(advec2);
  T maxdiffus = (T)P.nu * (T)P.dxyz_2;                                 // viscosity.f90:2435
  if constexpr (MAG) maxdiffus = fmax(maxdiffus, (T)P.eta * (T)P.dxyz_2);
  // diffus_chi: K-profile chix*dxyz_2 (entropy.f90:6094) + chi-const (:4907)
  if constexpr (ENT && LAY) maxdiffus = fmax(maxdiffus, dchi + ((T)P.gamma * (T)P.chi) * (T)P.dxyz_2);
  else if constexpr (ENT) maxdiffus = fmax(maxdiffus, ((T)P.gamma * (T)P.chi) * (T)P.dxyz_2);
  const T dt1_advec = maxadvec / (T)P.cdt;
  const T dt1_diffus = maxdiffus / (T)P.cdtv;
  return (double)psqrt<T>(dt1_advec * dt1_advec + dt1_diffus * dt1_diffus);
}

// ---------------------------------------------------------------------------
// The RHS is split in two halves so that a kernel may evaluate them on
// different threads (the field half and the flow half exchange four numbers
// per cell): field_point covers src/magnetic.f90 (B, J, J x B / rho,
// induction, Alfven speed), flow_point_pre and flow_point_post the rest.
// rhs_point runs the three in sequence; every kernel therefore executes the
// same arithmetic in the same order.

// What the field half hands to the flow half.
template <typename T>
struct FieldOut {
  T jxbr[3];    // J x B / rho                         magnetic.f90:4871-4888
  T j2;         // J^2 (ohmic heating, entropy runs)   magnetic.f90:6143
  T advec_va2;  // Alfven Courant term                 magnetic.f90:5180
};

// Pencils the flow half keeps between its two parts.
template <typename T>
struct FlowKeep {
  T rho1, TT1, cs2, advec_uu;
  T dchi;  // K-profile diffusion chix*dxyz_2 (LAY only)
};

// Magnetic pencils (magnetic.f90:4279-5225; sub.f90:2945-3113) and daa_dt
// (magnetic.f90:6247,6524).  d.a*, d.aij, d.d2a, d.mxa and d.u are read;
// wa[0..2] (already alpha-scaled) receive the induction RHS.
template <typename T, bool ENT, bool DT>
__device__ __forceinline__ void field_point(const Derivs<T>& d, T rho1, const Phys& P, double* wa,
                                            FieldOut<T>& o, double* b2_out) {
  const T u0 = d.u[0], u1 = d.u[1], u2v = d.u[2];
  T bb[3];
  bb[0] = d.aij[2][1] - d.aij[1][2];                                  // curl_mn
  bb[1] = d.aij[0][2] - d.aij[2][0];
  bb[2] = d.aij[1][0] - d.aij[0][1];
  const T b2 = bb[0] * bb[0] + bb[1] * bb[1] + bb[2] * bb[2];
  T uxb[3];
  uxb[0] = u1 * bb[2] - u2v * bb[1];                                  // cross_mn
  uxb[1] = u2v * bb[0] - u0 * bb[2];
  uxb[2] = u0 * bb[1] - u1 * bb[0];
  // curlb = curl_mn(bij) with bij(i,j) = eps_ikl d2A(k,j,l)
  T curlb[3];
  curlb[0] = (d.mxa[0][1] - d.d2a[0][1]) - (d.d2a[0][2] - d.mxa[1][2]);
  curlb[1] = (d.mxa[2][2] - d.d2a[1][2]) - (d.d2a[1][0] - d.mxa[0][0]);
  curlb[2] = (d.mxa[1][0] - d.d2a[2][0]) - (d.d2a[2][1] - d.mxa[2][1]);
  T del2a[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) del2a[i] = d.d2a[i][0] + d.d2a[i][1] + d.d2a[i][2];
  T jj[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) jj[i] = (T)P.mu01 * curlb[i];
  o.j2 = 0;
  if constexpr (ENT) o.j2 = jj[0] * jj[0] + jj[1] * jj[1] + jj[2] * jj[2];
  T jxb[3];
  jxb[0] = jj[1] * bb[2] - jj[2] * bb[1];
  jxb[1] = jj[2] * bb[0] - jj[0] * bb[2];
  jxb[2] = jj[0] * bb[1] - jj[1] * bb[0];
#pragma unroll
  for (int i = 0; i < 3; ++i) o.jxbr[i] = rho1 * jxb[i];
  o.advec_va2 = 0;
  if constexpr (DT) o.advec_va2 = advec_va2_of<T>(bb, rho1, P);
#pragma unroll
  for (int j = 0; j < 3; ++j) wa[j] = wa[j] + (double)(uxb[j] + (T)P.eta * del2a[j]);  // :6247,6524
  *b2_out = (double)b2;
}

// Hydro, density, EOS, energy and viscosity pencils and their RHS terms up to
// (not including) the magnetic contributions.  wf[0..2] uu, wf[3] lnrho,
// wf[4] ss (ENT), already alpha-scaled.  rho1 = exp(-lnrho) is taken as an
// argument so that both halves use the same value.  Without entropy the EOS
// is isothermal (gamma = 1) or, with POLY, polytropic (noentropy.f90:94-103);
// it is a compile-time choice so that no branch splits the algebra.
template <typename T, bool ENT, bool DT, bool POLY, int UPW = 0, bool LAY = false>
__device__ __forceinline__ void flow_point_pre(const Derivs<T>& d, T rho1, const Phys& P, double* wf,
                                               FlowKeep<T>& k, double* u2_out, const double* zl = nullptr) {
  // ---- hydro pencils (hydro.f90:3925-4000; sub.f90 helpers) ----
  const T u0 = d.u[0], u1 = d.u[1], u2v = d.u[2];
  const T u2 = u0 * u0 + u1 * u1 + u2v * u2v;                           // dot2_mn
  const T divu = d.uij[0][0] + d.uij[1][1] + d.uij[2][2];               // div_mn
  T s[3][3];                                                            // traceless_strain
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    s[j][j] = d.uij[j][j] - (T)(1.0 / 3.0) * divu;
#pragma unroll
    for (int i = j + 1; i < 3; ++i) {
      const T v = (T)0.5 * (d.uij[i][j] + d.uij[j][i]);
      s[i][j] = v;
      s[j][i] = v;
    }
  }
  T ugu[3];                                                             // u_dot_grad_vec
#pragma unroll
  for (int j = 0; j < 3; ++j) ugu[j] = u0 * d.uij[j][0] + u1 * d.uij[j][1] + u2v * d.uij[j][2];
  if constexpr ((UPW & UPW_UU) != 0) {  // doupwind: ugradf = ugradf - del6f_upwind
#pragma unroll
    for (int j = 0; j < 3; ++j) ugu[j] = ugu[j] - d.upw[j];
  }
  const T* del2u = d.del2u;                                             // del2v_etc
  const T* gdivu = d.gdivu;

  // ---- density pencils (density.f90:2727-2811) ----
  const T lnrho = d.lnrho;
  T uglnrho = u0 * d.glnrho[0] + u1 * d.glnrho[1] + u2v * d.glnrho[2];
  if constexpr ((UPW & UPW_LNRHO) != 0) uglnrho = uglnrho - d.upw[3];
  T sglnrho[3];                                                         // multmv_mn
#pragma unroll
  for (int i = 0; i < 3; ++i)
    sglnrho[i] = s[i][0] * d.glnrho[0] + s[i][1] * d.glnrho[1] + s[i][2] * d.glnrho[2];

  // ---- EOS + energy pencils ----
  T cs2, TT1 = 0, glnTT[3] = {0, 0, 0}, del2lnTT = 0, ugss = 0;
  T fpres[3];
  if constexpr (!ENT) {
    cs2 = eos_cs2<T, false, POLY>(lnrho, (T)0, P);
#pragma unroll
    for (int j = 0; j < 3; ++j) fpres[j] = -cs2 * d.glnrho[j];          // noentropy.f90:252-300
  } else {
    const T ss = d.ss;
    const T del2lnrho = d.del2lnrho;
    const T del2ss = d.del2ss;
    const T dl = lnrho - (T)P.lnrho0;
    cs2 = eos_cs2<T, true, false>(lnrho, ss, P);
    const T lnTT = (T)P.lnTT0 + (T)P.cv1 * ss + (T)P.gamma_m1 * dl;      // :876
    TT1 = pexp<T>(-lnTT);
#pragma unroll
    for (int j = 0; j < 3; ++j) glnTT[j] = (T)P.gamma_m1 * d.glnrho[j] + (T)P.cv1 * d.gss[j];
    del2lnTT = (T)P.gamma_m1 * del2lnrho + (T)P.cv1 * del2ss;
    ugss = u0 * d.gss[0] + u1 * d.gss[1] + u2v * d.gss[2];               // entropy.f90:3563
    if constexpr ((UPW & UPW_SS) != 0) ugss = ugss - d.upw[4];          // :3576
#pragma unroll
    for (int j = 0; j < 3; ++j)                                         // entropy.f90:3580
      fpres[j] = -(cs2 * (d.glnrho[j] + glnTT[j]) * (T)P.gamma1);
  }

  // ---- viscosity, nu-const (viscosity.f90:1633-1663) ----
  T fvisc[3];
#pragma unroll
  for (int j = 0; j < 3; ++j)
    fvisc[j] = (T)P.nu * (del2u[j] + (T)2. * sglnrho[j] + (T)(1. / 3.) * gdivu[j]);

  // ---- calc_all_rhs: duu_dt, dlnrho_dt, denergy_dt (equ.f90:1268-1274) ----
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    double acc = wf[j];
    acc = acc - (double)ugu[j];        // hydro.f90:4384
    if constexpr (LAY) {               // Coriolis force (hydro.f90:4494, :6339-6413)
      if (j == 0) acc = acc + (double)((T)P.cor_c2 * u1);
      else if (j == 1) acc = (acc - (double)((T)P.cor_c2 * u0)) + (double)((T)P.cor_s2 * u2v);
      else acc = acc - (double)((T)P.cor_s2 * u1);
    }
    acc = acc + (double)fvisc[j];      // viscosity.f90:2823 (calc_viscous_force, hydro.f90:4518)
    acc = acc + (double)fpres[j];      // noentropy.f90:381 / entropy.f90:3714
    wf[j] = acc;
  }
  {
    const T density_rhs = -divu - uglnrho;  // density.f90:2929-2930
    wf[3] = wf[3] + (double)density_rhs;
  }
  if constexpr (ENT) {
    const T visc_heat = (T)2 * (T)P.nu * d_sij2(s);
    double acc = wf[4];
    acc = acc - (double)ugss;                               // entropy.f90:3758
    acc = acc + (double)(TT1 * visc_heat);                  // viscosity.f90:2764
    if constexpr (LAY) {
      // K-profile conduction (entropy.f90:3800, :6046-6067): glnThcond =
      // glnTT + (0, 0, dlnhcond); thdiff = rho1*hcond*(del2lnTT + glnTT.glnThcond)
      if (P.kprof) {
        const T hc = (T)zl[0];
        const T gk = glnTT[0] * glnTT[0] + glnTT[1] * glnTT[1] + glnTT[2] * (glnTT[2] + (T)zl[1]);
        acc = acc + (double)(rho1 * hc * (del2lnTT + gk));
      }
    }
    const T a0 = d.glnrho[0] + glnTT[0], a1 = d.glnrho[1] + glnTT[1], a2 = d.glnrho[2] + glnTT[2];
    const T g2 = a0 * glnTT[0] + a1 * glnTT[1] + a2 * glnTT[2];
    acc = acc + (double)((T)P.chi * (T)P.cp * (del2lnTT + g2));  // entropy.f90:4852-4947
    if constexpr (LAY) {
      // heating and cooling (calc_heat_cool :3839, :6554; get_heat_cool_gravz
      // :6761-6828): heat = profz_heat - profz_cool*(cs2-cs2cool)/cs2cool
      if (P.heatcool) {
        const T heat = (T)zl[2] - (T)zl[3] * (cs2 - (T)P.cs2cool) / (T)P.cs2cool;
        acc = acc + (double)(TT1 * rho1 * heat);
      }
    }
    wf[4] = acc;
  }
  k.dchi = 0;
  if constexpr (LAY && ENT && DT) {
    if (P.kprof) k.dchi = rho1 * (T)zl[0] * (T)P.cv1 * (T)P.dxyz_2;  // chix*dxyz_2 (entropy.f90:6094)
  }
  k.rho1 = rho1;
  k.TT1 = TT1;
  k.cs2 = cs2;
  k.advec_uu = 0;
  if constexpr (DT) k.advec_uu = advec_uu_of<T>(u0, u1, u2v, P);
  *u2_out = (double)u2;
}

// The magnetic contributions to the flow RHS (Lorentz force last in duu_dt,
// magnetic.f90:5531; ohmic heating last in the entropy equation,
// magnetic.f90:6143) and set_dt1_max (equ.f90:1858-1912).
template <typename T, bool MAG, bool ENT, bool DT, bool LAY = false>
__device__ __forceinline__ void flow_point_post(const FlowKeep<T>& k, const FieldOut<T>& fo, const Phys& P,
                                                double* wf, double* dt1) {
  if constexpr (MAG) {
#pragma unroll
    for (int j = 0; j < 3; ++j) wf[j] = wf[j] + (double)fo.jxbr[j];
    if constexpr (ENT) wf[4] = wf[4] + (double)((T)P.eta * (T)P.mu0 * fo.j2 * k.rho1 * k.TT1);
  }
  // addgravity after the magnetic terms (equ.f90:1316; gravity_simple.f90:1083-1162);
  // adding 0 without gravity leaves the value unchanged
  wf[2] = wf[2] + P.gravz;
  if constexpr (DT) *dt1 = courant_dt1<T, MAG, ENT, LAY>(k.advec_uu, k.cs2, fo.advec_va2, P, k.dchi);
}

// Whole RHS of one cell: adds it to w[] (already alpha-scaled; slots 0..2 uu,
// 3 lnrho, [iss], [iax..iax+2]) and returns dt1_max_loc (when DT), u2 and b2.
template <typename T, bool MAG, bool ENT, bool DT, int UPW = 0, bool LAY = false>
__device__ __forceinline__ void rhs_point(const Derivs<T>& d, const Phys& P, double* w, int iss,
                                          int iax, double* dt1, double* u2_out, double* b2_out,
                                          const double* zl = nullptr) {
  const T rho1 = pexp<T>(-d.lnrho);
  FieldOut<T> fo{};
  *b2_out = 0.0;
  if constexpr (MAG) field_point<T, ENT, DT>(d, rho1, P, w + iax, fo, b2_out);
  double wf[5];
#pragma unroll
  for (int v = 0; v < 4; ++v) wf[v] = w[v];
  if constexpr (ENT) wf[4] = w[iss];
  FlowKeep<T> k;
  if (!ENT && P.polytropic) flow_point_pre<T, ENT, DT, true, UPW, LAY>(d, rho1, P, wf, k, u2_out, zl);
  else flow_point_pre<T, ENT, DT, false, UPW, LAY>(d, rho1, P, wf, k, u2_out, zl);
  flow_point_post<T, MAG, ENT, DT, LAY>(k, fo, P, wf, dt1);
#pragma unroll
  for (int v = 0; v < 4; ++v) w[v] = wf[v];
  if constexpr (ENT) w[iss] = wf[4];
}

}  // namespace pcmx
