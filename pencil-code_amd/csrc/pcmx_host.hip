// pcmx_host.hip -- host driver and C ABI of libpcmx (declared in include/pcmx.h).
//
// One engine per process, like the reference's `module GPU` singleton
// (src/gpu_astaroth.f90).  Device state: two ping-pong copies of the evolved
// variables (the fused kernel reads the stencil of f while writing the
// update, so it cannot update in place) and the 2N register w, each stored
// interior-only as [var][z][y][x] doubles; periodic ghosts are formed by the
// kernels.  The 2N-RK3 loop of timestep.f90:77-204 runs here, one fused
// kernel per substep.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cerrno>
#include <cstdint>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include "../../include/pcmx.h"
#include "pcmx_aux.h"
#include "pcmx_hostutil.h"
#include "pcmx_launch.h"

using namespace pcmx;

namespace pcmx {

thread_local char g_err[1024] = "";

int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

const char* last_error() { return g_err; }

// Williamson (1980) 2N-RK3, timestep.f90:43-44.
const double ALPHA_TS[3] = {0.0, -5.0 / 9.0, -153.0 / 128.0};
const double BETA_TS[3] = {1.0 / 3.0, 15.0 / 16.0, 8.0 / 15.0};


int check_params(const pcmx_params* p) {
  if (!p) return fail(PCMX_ERR_ARG, "params is NULL");
  if (p->nx < 1 || p->ny < 1 || p->nz < 1) return fail(PCMX_ERR_ARG, "grid %dx%dx%d", p->nx, p->ny, p->nz);
  if (p->nx < 3 || p->ny < 3 || p->nz < 3)
    return fail(PCMX_ERR_ARG, "periodic 6th-order stencils need >= 3 interior points per direction");
  const bool ent = p->iss >= 0, mag = p->iax >= 0;
  const int expect = 4 + (ent ? 1 : 0) + (mag ? 3 : 0);
  if (p->mvar != expect) return fail(PCMX_ERR_ARG, "mvar=%d but slots imply %d", p->mvar, expect);
  // registration order (register.f90:137-146): uu, lnrho, [ss], [aa]
  if (p->iux != 0 || p->ilnrho != 3 || (ent && p->iss != 4) || (mag && p->iax != (ent ? 5 : 4)))
    return fail(PCMX_ERR_ARG, "variable slots must follow the registration order uu,lnrho,[ss],[aa]");
  if (!(p->dx > 0 && p->dy > 0 && p->dz > 0)) return fail(PCMX_ERR_ARG, "grid spacing must be positive");
  if (p->precision != PCMX_PREC_FP64 && p->precision != PCMX_PREC_MIXED)
    return fail(PCMX_ERR_ARG, "unknown precision %d", p->precision);
  if (ent && p->gamma == 1.0) return fail(PCMX_ERR_ARG, "entropy needs gamma /= 1");
  if (p->cdt <= 0 || p->cdtv <= 0) return fail(PCMX_ERR_ARG, "cdt and cdtv must be positive");
  if (p->lperi_z != 0 && p->lperi_z != 1) return fail(PCMX_ERR_ARG, "lperi_z is 0 or 1");
  if (!std::isfinite(p->gravz)) return fail(PCMX_ERR_ARG, "gravz is not finite");
  if (!p->lperi_z) {
    if (p->nz < 8) return fail(PCMX_ERR_ARG, "a non-periodic z needs nz >= 8 (nz=%d)", p->nz);
    if (p->kernel == 2) return fail(PCMX_ERR_ARG, "kernel 2 runs periodic z only");
    for (int v = 0; v < p->mvar; ++v)
      for (int top = 0; top < 2; ++top) {
        const int c = top ? p->bcz_top[v] : p->bcz_bot[v];
        if (c == PCMX_BC_S || c == PCMX_BC_A || c == PCMX_BC_A2) continue;
        if (c != PCMX_BC_C1 && c != PCMX_BC_CT)
          return fail(PCMX_ERR_ARG, "bcz of variable %d: code %d is not a PCMX_BC_* code", v, c);
        if (v != p->iss) return fail(PCMX_ERR_ARG, "bcz 'c1' and 'cT' apply to the entropy slot only");
        if (c == PCMX_BC_C1 && !p->lheatc_kprof)
          return fail(PCMX_ERR_ARG, "bcz 'c1' needs the K-profile heat conduction (lheatc_kprof)");
        if (c == PCMX_BC_CT && !((top ? p->cs2top : p->cs2bot) > 0))
          return fail(PCMX_ERR_ARG, "bcz 'cT' needs cs2%s > 0", top ? "top" : "bot");
      }
  }
  // layer physics (samples/conv-slab)
  if (!std::isfinite(p->omega) || !std::isfinite(p->theta)) return fail(PCMX_ERR_ARG, "Omega or theta not finite");
  if ((p->lheatc_kprof != 0 && p->lheatc_kprof != 1) || (p->cooling_profile != 0 && p->cooling_profile != 1))
    return fail(PCMX_ERR_ARG, "lheatc_kprof is 0 or 1, cooling_profile 0 ('gaussian') or 1 ('step')");
  const bool heatcool = p->luminosity != 0.0 || p->cool != 0.0;
  if ((p->lheatc_kprof || heatcool) && !ent)
    return fail(PCMX_ERR_ARG, "K-profile conduction, heating and cooling need the entropy variable");
  if (p->lheatc_kprof && !(p->widthss != 0.0 && std::isfinite(p->hcond0)))
    return fail(PCMX_ERR_ARG, "K-profile conduction needs widthss /= 0");
  if (heatcool && !(p->cs2cool > 0)) return fail(PCMX_ERR_ARG, "heating or cooling needs cs2cool > 0");
  if (p->luminosity != 0.0 && !(p->wheat > 0)) return fail(PCMX_ERR_ARG, "heating needs wheat > 0");
  if (p->cool != 0.0 && !(p->wcool > 0)) return fail(PCMX_ERR_ARG, "cooling needs wcool > 0");
  if (layers_on(*p) && !(p->lxyz[0] > 0 && p->lxyz[1] > 0 && p->lxyz[2] > 0))
    return fail(PCMX_ERR_ARG, "the layer physics needs the box size lxyz");
  if (p->kernel < 0 || p->kernel > 3) return fail(PCMX_ERR_ARG, "unknown kernel %d", p->kernel);
  if ((p->lupw_uu != 0 && p->lupw_uu != 1) || (p->lupw_lnrho != 0 && p->lupw_lnrho != 1) ||
      (p->lupw_ss != 0 && p->lupw_ss != 1))
    return fail(PCMX_ERR_ARG, "lupw_uu, lupw_lnrho and lupw_ss are 0 or 1");
  if (p->lupw_ss && !ent) return fail(PCMX_ERR_ARG, "lupw_ss needs the entropy variable");
  const bool upw = p->lupw_uu || p->lupw_lnrho || p->lupw_ss;
  if (upw && p->kernel == 2) return fail(PCMX_ERR_ARG, "upwinding is not provided by kernel 2");
  const int mask = (p->lupw_uu ? UPW_UU : 0) | (p->lupw_lnrho ? UPW_LNRHO : 0) | (p->lupw_ss ? UPW_SS : 0);
  if (p->kernel == 3 && !split_upw_ok(ent, mask, TILE_32x8))
    return fail(PCMX_ERR_ARG, "kernel 3 upwinds uu+lnrho (isothermal), lnrho+ss or uu+lnrho+ss (entropy); "
                "use kernel 1 for other upwinding flags");
  if (layers_on(*p)) {
    if (p->kernel == 2 || (p->kernel == 3 && !split_lay_ok(ent, mask, TILE_32x8)))
      return fail(PCMX_ERR_ARG, "Omega, K-profile conduction, heating and cooling run on kernel 1, or on "
                  "kernel 3 with entropy, lnrho+ss upwinding and 32x8 tiles");
    if (!lay_upw_ok(ent, mask))
      return fail(PCMX_ERR_ARG, "with the layer physics the upwinding flags are none or lnrho, lnrho+ss, "
                  "uu+lnrho+ss (entropy), uu+lnrho (isothermal)");
  }
  return PCMX_OK;
}

// Per z plane {hcond, dlnhcond/dz, profz_heat, profz_cool} of the layer
// physics: get_gravz_heatcond (entropy.f90:7741-7769) with step / der_step
// (sub.f90:4975-5015, tini = 5*tiny), profz_heat (:1374-1386), profz_cool
// (:1319-1368) times cool.  z(n) = z0 + n*dz (non-periodic z) or
// z0 + (n+1/2)*dz.
void layer_row(const pcmx_params& p, int n, double out[4]) {
  const double tini = 5 * 2.2250738585072014e-308;
  auto step = [&](double x, double x0, double w) { return 0.5 * (1 + std::tanh((x - x0) / (w + tini))); };
  auto der_step = [&](double x, double x0, double w) {
    const double arg = std::fabs((x - x0) / (w + tini));
    if (std::fabs(arg) >= 8.) return 2. / w * std::exp(-2. * std::fabs(arg));
    const double c = std::cosh(arg);
    return 0.5 / (w * (c * c));
  };
  const double z = p.z0 + (p.lperi_z ? n + 0.5 : (double)n) * p.dz;
  out[0] = out[1] = out[2] = out[3] = 0.0;
  if (p.lheatc_kprof) {
    const double hc = 1. + (p.hcond1 - 1.) * step(z, p.z1, -p.widthss) + (p.hcond2 - 1.) * step(z, p.z2, p.widthss);
    out[1] = ((p.hcond1 - 1.) * der_step(z, p.z1, -p.widthss) + (p.hcond2 - 1.) * der_step(z, p.z2, p.widthss)) / hc;
    out[0] = p.hcond0 * hc;
  }
  if (p.luminosity != 0.0) {
    const double pi = 3.14159265358979323846264338327950;
    const double t = (z - p.z0) / p.wheat;
    double h = p.luminosity * std::exp(-0.5 * (t * t)) / (std::sqrt(pi / 2.) * p.wheat);
    if (p.ny == 1 && p.nx > 1) h = h / p.lxyz[0];
    else if (p.nx == 1 && p.ny > 1) h = h / p.lxyz[1];
    else if (p.nx > 1 && p.ny > 1) h = h / (p.lxyz[0] * p.lxyz[1]);
    out[2] = h;
  }
  if (p.cool != 0.0) {
    double c;
    if (p.cooling_profile == 1) {
      c = step(z, p.z2, p.wcool);
    } else {
      const double t = ((p.z0 + p.lxyz[2]) - z) / p.wcool;
      c = std::exp(-0.5 * (t * t));
    }
    out[3] = p.cool * c;
  }
}

// Constants of the 'c1' and 'cT' entropy rules of k_zbc (bc_ss_flux,
// bc_ss_temp_z; initialize_energy :851-873 for FbotKbot / FtopKtop;
// dz2_bound, grid.f90:3149-3151).
void zbc_entropy_constants(const pcmx_params& p, ZBC& b) {
  b.iss = p.iss;
  b.cs20 = p.cs0 * p.cs0;
  b.gamma_m1 = p.gamma - 1.0;
  b.cv1 = p.gamma * (1 / p.cp);
  b.lnrho0 = std::log(p.rho0);
  b.cp = p.cp;
  const double cv = (1 / p.gamma) * p.cp;
  b.cpmcv = p.cp - cv;
  const double kb = p.hcond0 * p.hcond1, kt = p.hcond0 * p.hcond2;
  b.fk[0] = kb != 0.0 ? p.fbot / kb : 0.0;
  b.fk[1] = kt != 0.0 ? p.ftop / kt : 0.0;
  b.tmpT[0] = p.cs2bot > 0 ? 2 * cv * std::log(p.cs2bot / b.cs20) : 0.0;
  b.tmpT[1] = p.cs2top > 0 ? 2 * cv * std::log(p.cs2top / b.cs20) : 0.0;
  const int n2 = p.nz - 1;
  for (int i = 1; i <= 3; ++i) {
    b.dz2[0][i - 1] = 2. * ((p.z0 + i * p.dz) - (p.z0 + 0 * p.dz));
    b.dz2[1][i - 1] = 2. * ((p.z0 + n2 * p.dz) - (p.z0 + (n2 - i) * p.dz));
  }
}

// Derived constants, computed in the reference's operation order
// (grid.f90:1045-1068, deriv.f90, eos_idealgas.f90:159-268, magnetic.f90:1553).
Phys make_phys(const pcmx_params& p) {
  Phys P{};
  P.dx_1 = 1.0 / p.dx;
  P.dy_1 = 1.0 / p.dy;
  P.dz_1 = 1.0 / p.dz;
  P.dxyz_2 = P.dx_1 * P.dx_1 + P.dy_1 * P.dy_1 + P.dz_1 * P.dz_1;
  const double a = 1.0 / 60.0;
  P.fac1[0] = a * P.dx_1;
  P.fac1[1] = a * P.dy_1;
  P.fac1[2] = a * P.dz_1;
  P.fac2[0] = P.dx_1 * P.dx_1;
  P.fac2[1] = P.dy_1 * P.dy_1;
  P.fac2[2] = P.dz_1 * P.dz_1;
  P.facxy = (1.0 / 720.0) * P.dx_1 * P.dy_1;
  P.facxz = (1.0 / 720.0) * P.dz_1 * P.dx_1;
  P.facyz = (1.0 / 720.0) * P.dy_1 * P.dz_1;
  P.nu = p.nu;
  P.eta = p.eta;
  P.chi = p.chi;
  P.cp = p.cp;
  P.mu0 = p.mu0;
  P.mu01 = 1. / p.mu0;
  P.gamma = p.gamma;
  P.gamma_m1 = p.gamma - 1.0;
  P.gamma1 = 1 / p.gamma;
  P.lnrho0 = std::log(p.rho0);
  P.cs20 = p.cs0 * p.cs0;
  const double cp1 = 1 / p.cp;
  P.cv1 = p.gamma * cp1;
  P.lnTT0 = (P.gamma_m1 != 0.0) ? std::log(P.cs20 / (p.cp * P.gamma_m1)) : std::log(P.cs20 / p.cp);
  P.cdt = p.cdt;
  P.cdtv = p.cdtv;
  P.polytropic = (p.iss < 0 && p.gamma != 1.0) ? 1 : 0;
  P.gravz = p.gravz;
  // coriolis_cartesian (hydro.f90:6339-6413); theta = 0 gives s2 = 0, which
  // the general form adds as +0
  if (p.theta == 0.0) {
    P.cor_c2 = 2 * p.omega;
    P.cor_s2 = 0.0;
  } else {
    const double dtor = 3.14159265358979323846264338327950 / 180.;
    P.cor_c2 = 2 * p.omega * std::cos(p.theta * dtor);
    P.cor_s2 = -2 * p.omega * std::sin(p.theta * dtor);
  }
  P.cs2cool = p.cs2cool;
  P.kprof = p.lheatc_kprof;
  P.heatcool = (p.luminosity != 0.0 || p.cool != 0.0) ? 1 : 0;
  P.zlay = nullptr;  // the engine's device table, set by the caller
  return P;
}

}  // namespace pcmx

namespace {

struct Engine {
  pcmx_params par{};
  Phys phys{};
  long nxyz = 0;
  int nvar = 0;
  bool mag = false, ent = false;
  double* f[2] = {nullptr, nullptr};
  double* w = nullptr;
  double* w2 = nullptr;       // second 2N register of the deferred Courant step (kernel 3)
  int cur = 0;
  double* partials = nullptr;
  long partials_cap = 0;
  double* results = nullptr;  // R_COUNT doubles
  double* staging = nullptr;  // one variable with ghosts (mw doubles)
  hipStream_t stream = nullptr;
  int kernel = 1;             // resolved kernel variant
  int tile = TILE_32x8;       // tile of the split kernel
  ZPlan zplan{};
  // timing of the fused kernel (pcmx_time_step)
  bool timing = false;
  double timing_ms = 0.0;
  long timing_n = 0;
  double last_dt = 0.0;
  // kernel 3 defers the first substep's update to the second substep: after
  // substep 1, f[cur] still holds f0 and f1 = f0 + dt*beta1*w1 is implicit
  bool pending_f1 = false;
  // pinned chunks of the device -> file snapshot path
  double* pin[2] = {nullptr, nullptr};
  hipEvent_t pin_ev[2] = {nullptr, nullptr};
  // helical forcing added after each step (pcmx_set_forcing)
  pcmx_forcing frc{};
  double2* ftab = nullptr;    // nx + ny + nz unit phasors
  // non-periodic z: ghost planes of the state ([v][6][ny][nx]) and the rules
  double* zgh = nullptr;
  ZBC zbc{};
  // layer physics: per z plane {hcond, dlnhcond, profz_heat, profz_cool}
  double* zlay = nullptr;
  // two fixed-dt steps captured as a HIP graph per [ldiag][cur], valid for
  // the dt and the parameter generation they were captured with
  hipGraphExec_t gexec[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
  double gdt[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
  long ggen[2][2] = {{-1, -1}, {-1, -1}};
  long gen = 0;               // bumped by every change of plan, physics or forcing
};

Engine* g = nullptr;

// ---------------------------------------------------------------- launches
long nblocks(const Engine& e) {
  return e.kernel >= 2 ? e.zplan.nblocks : direct_blocks(e.par.nx, e.par.ny, e.par.nz);
}

LaunchCfg cfg_of(const Engine& e, bool dt) {
  LaunchCfg c{};
  c.kernel = e.kernel;
  c.tile = e.tile;
  c.mag = e.mag;
  c.ent = e.ent;
  c.dt = dt;
  c.zp = e.zplan;
  c.upw = (e.par.lupw_uu ? UPW_UU : 0) | (e.par.lupw_lnrho ? UPW_LNRHO : 0) | (e.par.lupw_ss ? UPW_SS : 0);
  c.lay = layers_on(e.par);
  return c;
}

int launch_rhs(Engine& e, const KArgs& a, bool dt) {
  const LaunchCfg c = cfg_of(e, dt);
  const bool ok = (e.par.precision == PCMX_PREC_MIXED) ? launch_rhs_f32(c, a, e.phys, e.stream)
                                                       : launch_rhs_f64(c, a, e.phys, e.stream);
  if (!ok) return fail(PCMX_ERR_ARG, "no kernel variant for stage %d of kernel %d", a.stage, e.kernel);
  HIPCHECK(hipGetLastError());
  return PCMX_OK;
}

KArgs base_args(Engine& e) {
  KArgs a{};
  a.nx = e.par.nx;
  a.ny = e.par.ny;
  a.nz = e.par.nz;
  a.mvar = e.nvar;
  a.iss = e.par.iss;
  a.iax = e.par.iax;
  a.nxyz = e.nxyz;
  a.w = e.w;
  a.wout = e.w;
  a.means = e.results + R_MEAN0;
  a.zgh = e.zgh;
  return a;
}

// boundconds_z of a non-periodic z on state f: ghost planes into e.zgh and,
// with zero_a, the boundary planes of 'a' variables set to 0 (before an RHS,
// as the reference's pde does, equ.f90:209-224).  No-op for periodic z.
int apply_zbc(Engine& e, const double* f, bool zero_a) {
  if (!e.zgh) return PCMX_OK;
  const long nxy = (long)e.par.nx * e.par.ny;
  hipLaunchKernelGGL(k_zbc, dim3((unsigned)((nxy + 255) / 256), e.nvar), dim3(256), 0, e.stream,
                     const_cast<double*>(f), e.zgh, e.zbc, e.nvar, nxy, e.par.nz, zero_a ? 1 : 0);
  HIPCHECK(hipGetLastError());
  return PCMX_OK;
}

int finalize_red(Engine& e, bool set_dt, bool set_means) {
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, e.stream, e.partials, nblocks(e), e.results,
                     (double)e.nxyz, e.nvar, set_dt ? 1 : 0, set_means ? 1 : 0);
  HIPCHECK(hipGetLastError());
  return PCMX_OK;
}

// fact of the forcing of a step with time step dt: force/ffnorm*sqrt(dt)
// when the library applies the dt factor (forcing.f90:1958).
double forcing_fact(const Engine& e, double dt) {
  return e.frc.scale_sqrt_dt ? e.frc.fact * std::sqrt(dt) : e.frc.fact;
}

// The split kernel adds the forcing in the epilogue of the last substep
// (without entropy; with it, a separate pass follows the step).
bool forcing_fused(const Engine& e) { return e.kernel == 3 && !e.ent; }
void fuse_forcing(const Engine& e, KArgs& a, double dt) {
  if (!forcing_fused(e) || !e.frc.active) return;
  a.ftab = e.ftab;
  a.ffact = forcing_fact(e, dt);
  a.ftarget = e.frc.target;
}

// One RK substep s (0-based) on the device; no host synchronisation.
// dt_host is the fixed step (ignored when ldt).  diag requests the
// reductions of the first substep.
int substep(Engine& e, int s, double dt_host, bool diag, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr) {
  KArgs a = base_args(e);
  a.fin = e.f[e.cur];
  if (int rc = apply_zbc(e, a.fin, true)) return rc;
  a.fout = e.f[1 - e.cur];
  a.first = (s == 0);
  a.alpha = ALPHA_TS[s];
  a.beta = BETA_TS[s];
  a.dtb = dt_host * BETA_TS[s];  // dt_beta_ts = dt*beta_ts (timestep.f90:150,215)
  a.write_w = (s < 2);
  a.stage = (s == 0) ? ST_S1 : (s == 1 ? ST_S2 : ST_S3);
  const bool ldt = e.par.ldt != 0;
  if (ldt && s == 0) {
    // dt follows from this substep's RHS (equ.f90:408) before f is updated:
    // store w1, reduce dt1, then f1 = f0 + dt*beta1*w1.
    a.mode = MODE_W;
    a.stage = ST_W;
    a.partials = e.partials;
    if (ev0) HIPCHECK(hipEventRecord(ev0, e.stream));
    int rc = launch_rhs(e, a, true);
    if (rc) return rc;
    if (ev1) HIPCHECK(hipEventRecord(ev1, e.stream));
    if ((rc = finalize_red(e, true, true))) return rc;
    const long n = e.nxyz * e.nvar;
    hipLaunchKernelGGL(k_axpy_dt, dim3(2048), dim3(256), 0, e.stream, e.f[e.cur], e.f[1 - e.cur], e.w,
                       e.results + R_DT, BETA_TS[0], n);
    HIPCHECK(hipGetLastError());
  } else {
    a.mode = MODE_UPDATE;
    a.d_dt = ldt ? e.results + R_DT : nullptr;
    a.partials = (s == 0 && diag) ? e.partials : nullptr;
    if (s == 2) fuse_forcing(e, a, dt_host);
    if (ev0) HIPCHECK(hipEventRecord(ev0, e.stream));
    int rc = launch_rhs(e, a, false);
    if (rc) return rc;
    if (ev1) HIPCHECK(hipEventRecord(ev1, e.stream));
    if (s == 0 && diag)
      if ((rc = finalize_red(e, false, true))) return rc;
  }
  e.cur = 1 - e.cur;
  return PCMX_OK;
}

// One RK substep of a Courant-limited step with the deferred first-substep
// update (kernel 3): substep 1 stores only w1 = RHS(f0) and reduces dt1 (dt
// is then on the device before substep 2 needs it); substep 2 forms
// f1 = f0 + dt*beta1*w1 while staging and stores f2 and w2 (into the second
// 2N buffer: neighbouring blocks still read w1 in their halos); substep 3
// stores f3.  The buffer holding f3 is the one that held f0.
int substep_deferred(Engine& e, int s, double dt_host, bool diag, hipEvent_t ev0, hipEvent_t ev1) {
  KArgs a = base_args(e);
  const bool ldt = e.par.ldt != 0;
  a.alpha = ALPHA_TS[s];
  a.beta = BETA_TS[s];
  a.dtb = dt_host * BETA_TS[s];  // dt_beta_ts = dt*beta_ts (timestep.f90:150,215)
  a.beta1 = BETA_TS[0];
  a.dtb1 = dt_host * BETA_TS[0];
  a.d_dt = ldt ? e.results + R_DT : nullptr;
  a.first = (s == 0);
  a.write_w = (s < 2);
  a.fin = e.f[e.cur];
  a.fout = (s == 0) ? nullptr : e.f[1 - e.cur];
  if (s == 0) {
    a.stage = ST_W;
    a.mode = MODE_W;
    a.partials = e.partials;
  } else if (s == 1) {
    a.stage = ST_S2D;
    a.mode = MODE_UPDATE;
    a.wout = e.w2;
  } else {
    a.stage = ST_S3;
    a.mode = MODE_UPDATE;
    a.w = e.w2;
    fuse_forcing(e, a, dt_host);
  }
  if (ev0) HIPCHECK(hipEventRecord(ev0, e.stream));
  int rc = launch_rhs(e, a, s == 0 && ldt);
  if (rc) return rc;
  if (ev1) HIPCHECK(hipEventRecord(ev1, e.stream));
  (void)diag;
  if (s == 0) {
    if ((rc = finalize_red(e, true, true))) return rc;
    e.pending_f1 = true;
  } else {
    e.pending_f1 = false;
    e.cur = 1 - e.cur;
  }
  return PCMX_OK;
}

// The forcing of the step just completed, added to the state (addforce after
// time_step, run.f90:575).  dt is the step's dt, known on the host by now.
int add_forcing(Engine& e, double dt) {
  const dim3 grid((e.par.nx + 63) / 64, (e.par.ny + 3) / 4, e.par.nz);
  hipLaunchKernelGGL(k_add_forcing, grid, dim3(256), 0, e.stream, e.f[e.cur], e.ftab, forcing_fact(e, dt),
                     e.frc.target, e.par.nx, e.par.ny, e.par.nz);
  HIPCHECK(hipGetLastError());
  return PCMX_OK;
}

int run_substep(Engine& e, int s, double dt_host, bool diag, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr) {
  // the deferred first-substep update forms f1 while staging, which the
  // boundary conditions of a non-periodic z would have to follow: with one,
  // Courant steps take the explicit update
  const int rc = (e.kernel == 3 && e.par.ldt && e.par.lperi_z) ? substep_deferred(e, s, dt_host, diag, ev0, ev1)
                                                               : substep(e, s, dt_host, diag, ev0, ev1);
  if (rc || s != 2 || !e.frc.active || forcing_fused(e)) return rc;
  return add_forcing(e, dt_host);
}

// The state as the reference would hold it: after substep 1 of kernel 3,
// f1 = f0 + dt*beta1*w1 is formed into the spare buffer.  Returns the
// buffer to read.
double* current_state(Engine& e) {
  if (!e.pending_f1) return e.f[e.cur];
  hipLaunchKernelGGL(k_axpy_dt, dim3(2048), dim3(256), 0, e.stream, e.f[e.cur], e.f[1 - e.cur], e.w,
                     e.results + R_DT, BETA_TS[0], e.nxyz * e.nvar);
  return e.f[1 - e.cur];
}

int require() { return g ? PCMX_OK : fail(PCMX_ERR_STATE, "pcmx_initialize has not been called"); }

void free_engine(Engine* e) {
  if (!e) return;
  for (double* p : {e->f[0], e->f[1], e->w, e->w2, e->partials, e->results, e->staging})
    if (p) (void)hipFree(p);
  if (e->ftab) (void)hipFree(e->ftab);
  if (e->zgh) (void)hipFree(e->zgh);
  if (e->zlay) (void)hipFree(e->zlay);
  for (auto& row : e->gexec)
    for (auto& x : row)
      if (x) (void)hipGraphExecDestroy(x);
  for (int k = 0; k < 2; ++k) {
    if (e->pin[k]) (void)hipHostFree(e->pin[k]);
    if (e->pin_ev[k]) (void)hipEventDestroy(e->pin_ev[k]);
  }
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
}

long mw_of(const pcmx_params& p) { return (long)(p.nx + 6) * (p.ny + 6) * (p.nz + 6); }

// One Fortran unformatted sequential record, written as the gfortran runtime
// writes it: a 4-byte byte count before and after the data; a record longer
// than max_sub bytes (2147483639 by default) is split into subrecords, each
// framed the same way, whose leading count is negated when another
// subrecord follows and whose trailing count is negated when one precedes.
// The total length is known up front; data may arrive in pieces of any size.
class FRecord {
 public:
  static constexpr uint64_t MAX_SUB = 2147483639ULL;
  FRecord(FILE* fp, uint64_t total, uint64_t max_sub) : fp_(fp), left_(total), max_sub_(max_sub ? max_sub : MAX_SUB) {
    begin();
  }
  bool put(const void* data, uint64_t n) {
    const char* c = static_cast<const char*>(data);
    while (n > 0 && ok_) {
      if (in_sub_ == sub_len_) {  // only when a piece ends exactly at a subrecord boundary
        end();
        begin();
      }
      const uint64_t k = std::min(n, sub_len_ - in_sub_);
      ok_ = ok_ && std::fwrite(c, 1, k, fp_) == k;
      in_sub_ += k;
      left_ -= k;
      c += k;
      n -= k;
    }
    return ok_;
  }
  bool finish() {  // closes the last subrecord; every byte must have arrived
    if (left_ != 0) return false;
    end();
    return ok_;
  }

 private:
  void marker(int64_t v) {
    const int32_t m = (int32_t)v;
    ok_ = ok_ && std::fwrite(&m, 4, 1, fp_) == 1;
  }
  void begin() {
    sub_len_ = std::min(left_, max_sub_);
    in_sub_ = 0;
    marker(left_ > sub_len_ ? -(int64_t)sub_len_ : (int64_t)sub_len_);
  }
  void end() {
    marker(continued_ ? -(int64_t)sub_len_ : (int64_t)sub_len_);
    continued_ = true;
  }
  FILE* fp_;
  uint64_t left_, max_sub_, sub_len_ = 0, in_sub_ = 0;
  bool continued_ = false, ok_ = true;
};

}  // namespace

// ================================================================ C ABI
extern "C" {

const char* pcmx_last_error(void) { return g_err; }

int pcmx_register(void) { return PCMX_OK; }

int pcmx_initialize(const double* f, int mfarray, const pcmx_params* par, double t, int nt) {
  (void)t;
  int rc = check_params(par);
  if (rc) return rc;
  if (g) {
    free_engine(g);
    g = nullptr;
  }
  Engine* e = new Engine();
  e->par = *par;
  e->phys = make_phys(*par);
  e->nxyz = (long)par->nx * par->ny * par->nz;
  e->nvar = par->mvar;
  e->mag = par->iax >= 0;
  e->ent = par->iss >= 0;
  // kernel choice: the z-marching LDS kernels need whole tiles in x and y
  int kern = par->kernel;
  const int tile = (par->nz >= 4) ? split_tile(par->nx, par->ny) : -1;
  // upwinding of uu alone, of lnrho alone, or with entropy: the direct kernel
  const int upw_mask = (par->lupw_uu ? UPW_UU : 0) | (par->lupw_lnrho ? UPW_LNRHO : 0) | (par->lupw_ss ? UPW_SS : 0);
  const bool upw_split = tile >= 0 && (layers_on(*par) ? split_lay_ok(par->iss >= 0, upw_mask, tile)
                                                        : split_upw_ok(par->iss >= 0, upw_mask, tile));
  if (kern == 0) kern = (tile >= 0 && upw_split) ? 3 : 1;
  if (kern == 2 && !zmarch_supported(*par)) {
    delete e;
    return fail(PCMX_ERR_ARG, "z-marching kernel needs nx %% %d == 0 and ny %% %d == 0", ZM_TX, ZM_TY);
  }
  if (kern == 3 && tile < 0) {
    delete e;
    return fail(PCMX_ERR_ARG, "z-marching kernel needs nx %% 32 == 0 and ny %% 8 == 0, or nx and ny multiples of 16");
  }
  if (kern == 3 && !upw_split) {
    delete e;
    return fail(PCMX_ERR_ARG, "kernel 3 upwinds with entropy, and runs the layer physics, on 32x8 tiles only "
                "(nx %% 32 == 0, ny %% 8 == 0); "
                "use kernel 1");
  }
  e->kernel = kern;
  if (kern == 2) e->zplan = zmarch_plan(*par);
  if (kern == 3) {
    const char* env = std::getenv("PCMX_ZM_LZ");
    e->tile = tile;
    e->zplan = tile_plan(par->nx, par->ny, par->nz, tile_tx(tile), tile_ty(tile), env ? std::atoi(env) : 0);
  }
  const size_t bytes = sizeof(double) * e->nxyz * e->nvar;
  auto alloc = [&](double** p, size_t n) -> int {
    hipError_t err = hipMalloc((void**)p, n);
    if (err != hipSuccess)
      return fail(PCMX_ERR_HIP, "hipMalloc(%zu bytes) failed: %s", n, hipGetErrorString(err));
    return PCMX_OK;
  };
  e->partials_cap = std::max(direct_blocks(par->nx, par->ny, par->nz), e->kernel >= 2 ? e->zplan.nblocks : 0L);
  if ((rc = alloc(&e->f[0], bytes)) || (rc = alloc(&e->f[1], bytes)) || (rc = alloc(&e->w, bytes)) ||
      (rc = alloc(&e->partials, sizeof(double) * NRED * e->partials_cap)) ||
      (rc = alloc(&e->results, sizeof(double) * R_COUNT)) ||
      (e->kernel == 3 && par->ldt && (rc = alloc(&e->w2, bytes))) ||
      (!par->lperi_z && (rc = alloc(&e->zgh, sizeof(double) * 6 * par->nx * par->ny * e->nvar)))) {
    free_engine(e);
    return rc;
  }
  for (int v = 0; v < 8; ++v) {
    e->zbc.bot[v] = par->bcz_bot[v];
    e->zbc.top[v] = par->bcz_top[v];
  }
  zbc_entropy_constants(*par, e->zbc);
  if (layers_on(*par)) {
    std::vector<double> tab(4 * (size_t)par->nz);
    for (int n = 0; n < par->nz; ++n) layer_row(*par, n, &tab[4 * (size_t)n]);
    if ((rc = alloc(&e->zlay, sizeof(double) * tab.size()))) {
      free_engine(e);
      return rc;
    }
    hipError_t cerr = hipMemcpy(e->zlay, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice);
    if (cerr != hipSuccess) {
      free_engine(e);
      return fail(PCMX_ERR_HIP, "hipMemcpy(layer table) failed: %s", hipGetErrorString(cerr));
    }
    e->phys.zlay = e->zlay;
  }
  hipError_t err = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
  if (err != hipSuccess) {
    free_engine(e);
    return fail(PCMX_ERR_HIP, "hipStreamCreate failed: %s", hipGetErrorString(err));
  }
  (void)hipMemsetAsync(e->w, 0, bytes, e->stream);
  (void)hipMemsetAsync(e->results, 0, sizeof(double) * R_COUNT, e->stream);
  g = e;
  if (nt > 0) {
    if (!f) return fail(PCMX_ERR_ARG, "nt > 0 needs the host f-array");
    if ((rc = pcmx_load_farray(f, mfarray))) return rc;
  }
  return PCMX_OK;
}

int pcmx_finalize(void) {
  if (g) {
    (void)hipStreamSynchronize(g->stream);
    free_engine(g);
    g = nullptr;
  }
  return PCMX_OK;
}

int pcmx_load_farray(const double* f, int mfarray) {
  int rc = require();
  if (rc) return rc;
  Engine& e = *g;
  if (!f) return fail(PCMX_ERR_ARG, "f is NULL");
  if (mfarray < e.nvar) return fail(PCMX_ERR_ARG, "mfarray=%d < mvar=%d", mfarray, e.nvar);
  const long mw = mw_of(e.par);
  if (!e.staging) HIPCHECK(hipMalloc((void**)&e.staging, sizeof(double) * mw));
  for (int v = 0; v < e.nvar; ++v) {
    HIPCHECK(hipMemcpyAsync(e.staging, f + (long)v * mw, sizeof(double) * mw, hipMemcpyHostToDevice, e.stream));
    hipLaunchKernelGGL(k_unpack_interior, dim3(1024), dim3(256), 0, e.stream, e.staging,
                       e.f[e.cur] + (long)v * e.nxyz, e.par.nx, e.par.ny, e.par.nz);
    HIPCHECK(hipGetLastError());
  }
  // box means for the mixed form, diagnostics and dt1 of the loaded state
  e.pending_f1 = false;
  KArgs a = base_args(e);
  a.fin = e.f[e.cur];
  a.fout = nullptr;
  a.first = 1;
  a.mode = MODE_DT;
  a.stage = ST_DT;
  a.partials = e.partials;
  // the mixed form needs means before its first use: compute them with the
  // FP64 variant of the engine's kernel
  const int saved_prec = e.par.precision;
  e.par.precision = PCMX_PREC_FP64;
  rc = launch_rhs(e, a, true);
  e.par.precision = saved_prec;
  if (rc) return rc;
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, e.stream, e.partials, nblocks(e), e.results,
                     (double)e.nxyz, e.nvar, 0, 1);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(e.stream));
  return PCMX_OK;
}

int pcmx_copy_farray(double* f, int mfarray) {
  int rc = require();
  if (rc) return rc;
  Engine& e = *g;
  if (!f) return fail(PCMX_ERR_ARG, "f is NULL");
  if (mfarray < e.nvar) return fail(PCMX_ERR_ARG, "mfarray=%d < mvar=%d", mfarray, e.nvar);
  const long mw = mw_of(e.par);
  if (!e.staging) HIPCHECK(hipMalloc((void**)&e.staging, sizeof(double) * mw));
  const double* state = current_state(e);
  if ((rc = apply_zbc(e, state, false))) return rc;  // z ghosts of the current state
  const long nxy = (long)e.par.nx * e.par.ny;
  for (int v = 0; v < e.nvar; ++v) {
    hipLaunchKernelGGL(k_pack_ghosts, dim3(1024), dim3(256), 0, e.stream, state + (long)v * e.nxyz,
                       e.staging, e.par.nx, e.par.ny, e.par.nz, e.zgh ? e.zgh + v * 6 * nxy : nullptr);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(f + (long)v * mw, e.staging, sizeof(double) * mw, hipMemcpyDeviceToHost, e.stream));
    HIPCHECK(hipStreamSynchronize(e.stream));
  }
  return PCMX_OK;
}

int pcmx_before_boundary(int lrmv, int isubstep, double t) {
  (void)lrmv;
  (void)isubstep;
  (void)t;
  return require();
}

int pcmx_rhs(int isubstep, double t, double* dt) {
  (void)t;
  int rc = require();
  if (rc) return rc;
  Engine& e = *g;
  if (isubstep < 1 || isubstep > 3) return fail(PCMX_ERR_ARG, "isubstep=%d (itorder=3)", isubstep);
  if (!dt) return fail(PCMX_ERR_ARG, "dt is NULL");
  const int s = isubstep - 1;
  if (!e.par.ldt) {
    e.last_dt = *dt;
    double r[1] = {*dt};
    HIPCHECK(hipMemcpyAsync(e.results + R_DT, r, sizeof(double), hipMemcpyHostToDevice, e.stream));
  }
  if ((rc = run_substep(e, s, *dt, true))) return rc;
  if (e.par.ldt && s == 0) {
    double r;
    HIPCHECK(hipMemcpyAsync(&r, e.results + R_DT, sizeof(double), hipMemcpyDeviceToHost, e.stream));
    HIPCHECK(hipStreamSynchronize(e.stream));
    *dt = r;
    e.last_dt = r;
  } else {
    HIPCHECK(hipStreamSynchronize(e.stream));
  }
  return PCMX_OK;
}

int pcmx_update_after_substep(void) { return require(); }
int pcmx_split_update(void) { return require(); }

int pcmx_prepare_for_first_substep(double t, double* dt1) {
  (void)t;
  int rc = require();
  if (rc) return rc;
  Engine& e = *g;
  KArgs a = base_args(e);
  a.fin = current_state(e);
  if ((rc = apply_zbc(e, a.fin, true))) return rc;
  a.first = 1;
  a.mode = MODE_DT;
  a.stage = ST_DT;
  a.partials = e.partials;
  if ((rc = launch_rhs(e, a, true))) return rc;
  if ((rc = finalize_red(e, false, false))) return rc;
  double r[R_COUNT];
  HIPCHECK(hipMemcpyAsync(r, e.results, sizeof(r), hipMemcpyDeviceToHost, e.stream));
  HIPCHECK(hipStreamSynchronize(e.stream));
  if (dt1) *dt1 = r[R_DT1];
  return PCMX_OK;
}

int pcmx_get_reduced_vars(double* dst) {
  int rc = require();
  if (rc) return rc;
  if (!dst) return fail(PCMX_ERR_ARG, "dst is NULL");
  Engine& e = *g;
  double r[R_COUNT];
  HIPCHECK(hipMemcpyAsync(r, e.results, sizeof(r), hipMemcpyDeviceToHost, e.stream));
  HIPCHECK(hipStreamSynchronize(e.stream));
  for (int i = 0; i < PCMX_N_REDUCED; ++i) dst[i] = 0.0;
  dst[0] = r[R_URMS];
  dst[1] = r[R_UMAX];
  dst[2] = r[R_BRMS];
  dst[3] = r[R_DT1];
  dst[4] = e.par.ldt ? r[R_DT] : e.last_dt;
  for (int v = 0; v < e.nvar; ++v) dst[5 + v] = r[R_MEAN0 + v];
  return PCMX_OK;
}

static const char* const kParNames[] = {"nu", "eta", "chi", "cs0", "cdt", "cdtv"};

int pcmx_update_by_index(int index, double value) {
  int rc = require();
  if (rc) return rc;
  Engine& e = *g;
  switch (index) {
    case 0: e.par.nu = value; break;
    case 1: e.par.eta = value; break;
    case 2: e.par.chi = value; break;
    case 3: e.par.cs0 = value; break;
    case 4: e.par.cdt = value; break;
    case 5: e.par.cdtv = value; break;
    default: return fail(PCMX_ERR_NOTFOUND, "no parameter with index %d", index);
  }
  if ((rc = check_params(&e.par))) return rc;
  e.phys = make_phys(e.par);
  e.phys.zlay = e.zlay;
  ++e.gen;
  return PCMX_OK;
}

int pcmx_update_by_name(const char* varname, double value) {
  if (!varname) return fail(PCMX_ERR_ARG, "varname is NULL");
  for (int i = 0; i < (int)(sizeof(kParNames) / sizeof(kParNames[0])); ++i)
    if (std::strcmp(varname, kParNames[i]) == 0) {
      int rc = pcmx_update_by_index(i, value);
      return rc ? rc : i;
    }
  return fail(PCMX_ERR_NOTFOUND, "variable \"%s\" not found", varname);
}

int pcmx_test_bcs(double* maxdiff) {
  int rc = require();
  if (rc) return rc;
  if (!maxdiff) return fail(PCMX_ERR_ARG, "maxdiff is NULL");
  Engine& e = *g;
  const pcmx_params& p = e.par;
  const long n = e.nxyz, mw = mw_of(p);
  std::vector<double> in(n), out(mw);
  std::mt19937_64 rng(12345);
  std::uniform_real_distribution<double> U(-1.0, 1.0);
  for (auto& v : in) v = U(rng);
  double* d_in = nullptr;
  double* d_out = nullptr;
  HIPCHECK(hipMalloc((void**)&d_in, sizeof(double) * n));
  if (hipMalloc((void**)&d_out, sizeof(double) * mw) != hipSuccess) {
    (void)hipFree(d_in);
    return fail(PCMX_ERR_HIP, "hipMalloc failed in pcmx_test_bcs");
  }
  (void)hipMemcpy(d_in, in.data(), sizeof(double) * n, hipMemcpyHostToDevice);
  hipLaunchKernelGGL(k_pack_ghosts, dim3(1024), dim3(256), 0, e.stream, d_in, d_out, p.nx, p.ny, p.nz, nullptr);
  (void)hipStreamSynchronize(e.stream);
  (void)hipMemcpy(out.data(), d_out, sizeof(double) * mw, hipMemcpyDeviceToHost);
  (void)hipFree(d_in);
  (void)hipFree(d_out);
  // host reference: boundconds_x, then y, then z over full extents
  const int mx = p.nx + 6, my = p.ny + 6, mz = p.nz + 6;
  std::vector<double> h(mw, 0.0);
  for (int z = 0; z < p.nz; ++z)
    for (int y = 0; y < p.ny; ++y)
      for (int x = 0; x < p.nx; ++x)
        h[((long)(z + 3) * my + y + 3) * mx + x + 3] = in[((long)z * p.ny + y) * p.nx + x];
  for (int iz = 0; iz < mz; ++iz)
    for (int iy = 0; iy < my; ++iy)
      for (int gg = 0; gg < 3; ++gg) {
        double* row = &h[((long)iz * my + iy) * mx];
        row[gg] = row[p.nx + gg];
        row[3 + p.nx + gg] = row[3 + gg];
      }
  for (int iz = 0; iz < mz; ++iz)
    for (int gg = 0; gg < 3; ++gg)
      for (int ix = 0; ix < mx; ++ix) {
        h[((long)iz * my + gg) * mx + ix] = h[((long)iz * my + p.ny + gg) * mx + ix];
        h[((long)iz * my + 3 + p.ny + gg) * mx + ix] = h[((long)iz * my + 3 + gg) * mx + ix];
      }
  for (int gg = 0; gg < 3; ++gg)
    for (long q = 0; q < (long)mx * my; ++q) {
      h[(long)gg * mx * my + q] = h[(long)(p.nz + gg) * mx * my + q];
      h[(long)(3 + p.nz + gg) * mx * my + q] = h[(long)(3 + gg) * mx * my + q];
    }
  double md = 0.0;
  for (long q = 0; q < mw; ++q) md = std::max(md, std::fabs(h[q] - out[q]));
  *maxdiff = md;
  return PCMX_OK;
}

int pcmx_tune(double* best_ms) {
  int rc = require();
  if (rc) return rc;
  Engine& e = *g;
  if (e.kernel != 3) return fail(PCMX_ERR_ARG, "pcmx_tune tunes the split z-marching kernel (kernel 3) only");
  // candidates: every tile shape the grid allows, chunk lengths dividing nz
  std::vector<std::pair<int, int>> cands;
  const int upw_mask = cfg_of(e, false).upw;
  for (int tile : {TILE_32x8, TILE_16x16}) {
    if (e.par.nx % tile_tx(tile) || e.par.ny % tile_ty(tile)) continue;
    const bool ok = layers_on(e.par) ? split_lay_ok(e.ent, upw_mask, tile) : split_upw_ok(e.ent, upw_mask, tile);
    if (!ok) continue;  // no variant for this tile
    for (int lz : {8, 16, 32, 64, 128, 256, 512, 1024})
      if (lz <= e.par.nz && e.par.nz % lz == 0) cands.push_back({tile, lz});
    if (e.par.nz > 1024 || e.par.nz % 8) cands.push_back({tile, e.par.nz});
  }
  // Time the three stages a step runs (fixed dt: S1, S2, S3; Courant steps:
  // W, S2D, S3), each reading the state f[cur] and writing only the spare
  // buffer and the 2N registers, which hold nothing between steps, so the
  // state is untouched.
  if (e.pending_f1) return fail(PCMX_ERR_STATE, "pcmx_tune between substeps 1 and 2 of a step");
  if ((rc = apply_zbc(e, e.f[e.cur], false))) return rc;
  const bool ldt = e.par.ldt != 0;
  const bool deferred = ldt && e.par.lperi_z;  // run_substep's choice
  KArgs st[3];
  for (int s = 0; s < 3; ++s) {
    KArgs& a = st[s];
    a = base_args(e);
    a.fin = e.f[e.cur];
    a.fout = e.f[1 - e.cur];
    a.alpha = ALPHA_TS[s];
    a.beta = BETA_TS[s];
    a.dtb = 0.0;
    a.beta1 = BETA_TS[0];
    a.dtb1 = 0.0;
    a.first = (s == 0);
    a.write_w = (s < 2);
    a.mode = MODE_UPDATE;
    a.stage = s == 0 ? ST_S1 : (s == 1 ? ST_S2 : ST_S3);
    a.d_dt = ldt ? e.results + R_DT : nullptr;
  }
  if (ldt) {
    st[0].stage = ST_W;
    st[0].mode = MODE_W;
    st[0].fout = nullptr;
  }
  if (deferred) {
    st[1].stage = ST_S2D;
    st[1].wout = e.w2;
    st[2].w = e.w2;
  }
  auto run_step = [&]() -> int {
    for (int s = 0; s < 3; ++s) {
      st[s].partials = (s == 0) ? e.partials : nullptr;
      const int r = launch_rhs(e, st[s], ldt && s == 0);
      if (r) return r;
    }
    return PCMX_OK;
  };
  const int tile0 = e.tile;
  const ZPlan zp0 = e.zplan;
  hipEvent_t t0, t1;
  HIPCHECK(hipEventCreate(&t0));
  HIPCHECK(hipEventCreate(&t1));
  float best = 1e30f;
  int btile = tile0;
  ZPlan bplan = zp0;
  for (const auto& c : cands) {
    e.tile = c.first;
    e.zplan = tile_plan(e.par.nx, e.par.ny, e.par.nz, tile_tx(c.first), tile_ty(c.first), c.second);
    if (e.zplan.nblocks > e.partials_cap) continue;
    if ((rc = run_step())) break;  // warm-up
    // two steps, or as many as fill about 20 ms so that short launches are
    // timed above the noise of the launch gaps
    HIPCHECK(hipEventRecord(t0, e.stream));
    if ((rc = run_step())) break;
    HIPCHECK(hipEventRecord(t1, e.stream));
    HIPCHECK(hipEventSynchronize(t1));
    float ms1 = 0.f;
    HIPCHECK(hipEventElapsedTime(&ms1, t0, t1));
    const int reps = std::max(2, std::min(50, (int)(20.0f / std::max(ms1, 1e-3f))));
    HIPCHECK(hipEventRecord(t0, e.stream));
    for (int it = 0; it < reps && !rc; ++it) rc = run_step();
    if (rc) break;
    HIPCHECK(hipEventRecord(t1, e.stream));
    HIPCHECK(hipEventSynchronize(t1));
    float ms = 0.f;
    HIPCHECK(hipEventElapsedTime(&ms, t0, t1));
    ms *= 2.0f / reps;  // per two steps
    if (ms < best) {
      best = ms;
      btile = c.first;
      bplan = e.zplan;
    }
  }
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  e.tile = rc ? tile0 : btile;
  e.zplan = rc ? zp0 : bplan;
  ++e.gen;
  if (rc) return rc;
  if (best_ms) {
    best_ms[0] = best / 6.0;  // per launch
    best_ms[1] = tile_tx(btile);
    best_ms[2] = bplan.lz;
  }
  return PCMX_OK;
}

int pcmx_set_timing(int enable) {
  int rc = require();
  if (rc) return rc;
  g->timing = enable != 0;
  return PCMX_OK;
}

int pcmx_kernel_timing(double* out) {
  int rc = require();
  if (rc) return rc;
  if (!out) return fail(PCMX_ERR_ARG, "out is NULL");
  out[0] = g->timing_n ? g->timing_ms / g->timing_n : 0.0;
  out[1] = (double)g->timing_n;
  return PCMX_OK;
}

int pcmx_time_step(int nsteps, double* t, double* dt, int ldiag) {
  int rc = require();
  if (rc) return rc;
  Engine& e = *g;
  if (nsteps < 0 || !t || !dt) return fail(PCMX_ERR_ARG, "bad arguments to pcmx_time_step");
  const bool ldt = e.par.ldt != 0;
  if (!ldt) {
    double r[1] = {*dt};
    HIPCHECK(hipMemcpyAsync(e.results + R_DT, r, sizeof(double), hipMemcpyHostToDevice, e.stream));
    e.last_dt = *dt;
  }
  std::vector<hipEvent_t> evs;
  if (e.timing) {
    evs.resize(6L * nsteps);
    for (auto& ev : evs) HIPCHECK(hipEventCreate(&ev));
  }
  // t advances as timestep.f90:108-149 does: ds = alpha*ds + 1, t += ds*dt*beta
  auto advance_t = [&]() {
    double ds = 0.0;
    for (int s = 0; s < 3; ++s) {
      ds = (s == 0) ? 1.0 : ALPHA_TS[s] * ds + 1.0;
      *t = *t + ds * (*dt * BETA_TS[s]);
    }
  };
  int it = 0;
  // Fixed steps without kernel timing replay a captured pair of steps (the
  // pair brings the state buffers back to their roles): one graph launch
  // instead of four or more kernel launches per step.
  if (!ldt && !e.timing) {
    const int di = ldiag ? 1 : 0;
    while (nsteps - it >= 2) {
      hipGraphExec_t& x = e.gexec[di][e.cur];
      if (x && (e.gdt[di][e.cur] != *dt || e.ggen[di][e.cur] != e.gen)) {
        (void)hipGraphExecDestroy(x);
        x = nullptr;
      }
      if (!x) {
        const int cur0 = e.cur;
        hipGraph_t graph = nullptr;
        HIPCHECK(hipStreamBeginCapture(e.stream, hipStreamCaptureModeThreadLocal));
        for (int k = 0; k < 6 && !rc; ++k) rc = run_substep(e, k % 3, *dt, ldiag != 0);
        const hipError_t ce = hipStreamEndCapture(e.stream, &graph);
        e.cur = cur0;
        if (rc || ce != hipSuccess) {
          if (graph) (void)hipGraphDestroy(graph);
          return rc ? rc : fail(PCMX_ERR_HIP, "hipStreamEndCapture failed: %s", hipGetErrorString(ce));
        }
        const hipError_t ie = hipGraphInstantiate(&x, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ie != hipSuccess) {
          x = nullptr;
          return fail(PCMX_ERR_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(ie));
        }
        e.gdt[di][e.cur] = *dt;
        e.ggen[di][e.cur] = e.gen;
      }
      HIPCHECK(hipGraphLaunch(x, e.stream));
      advance_t();
      advance_t();
      it += 2;
    }
  }
  for (; it < nsteps; ++it) {
    for (int s = 0; s < 3; ++s) {
      hipEvent_t e0 = e.timing ? evs[6L * it + 2 * s] : nullptr;
      hipEvent_t e1 = e.timing ? evs[6L * it + 2 * s + 1] : nullptr;
      if ((rc = run_substep(e, s, *dt, ldiag != 0, e0, e1))) return rc;
      if (ldt && s == 0) {
        // t advances with the step of this substep's Courant condition
        double r;
        HIPCHECK(hipMemcpyAsync(&r, e.results + R_DT, sizeof(double), hipMemcpyDeviceToHost, e.stream));
        HIPCHECK(hipStreamSynchronize(e.stream));
        *dt = r;
        e.last_dt = r;
      }
    }
    advance_t();
  }
  HIPCHECK(hipStreamSynchronize(e.stream));
  if (e.timing) {
    double tot = 0.0;
    for (int it = 0; it < nsteps; ++it)
      for (int s = 0; s < 3; ++s) {
        float ms = 0.f;
        HIPCHECK(hipEventElapsedTime(&ms, evs[6L * it + 2 * s], evs[6L * it + 2 * s + 1]));
        tot += ms;
      }
    e.timing_ms = tot;
    e.timing_n = 3L * nsteps;
    for (auto& ev : evs) (void)hipEventDestroy(ev);
  }
  return PCMX_OK;
}

// ------------------------------------------------------------------ forcing

int pcmx_set_forcing(const pcmx_forcing* frc) {
  int rc = require();
  if (rc) return rc;
  Engine& e = *g;
  if (!frc) return fail(PCMX_ERR_ARG, "frc is NULL");
  ++e.gen;
  if (!frc->active) {
    e.frc = pcmx_forcing{};
    return PCMX_OK;
  }
  const pcmx_params& p = e.par;
  if (frc->active != 1 || frc->reserved != 0 || (frc->scale_sqrt_dt != 0 && frc->scale_sqrt_dt != 1))
    return fail(PCMX_ERR_ARG, "pcmx_forcing: active and scale_sqrt_dt are 0 or 1, reserved is 0");
  if (frc->target != p.iux && !(e.mag && frc->target == p.iax))
    return fail(PCMX_ERR_ARG, "pcmx_forcing: target %d is neither iux (%d) nor iax (%d)", frc->target, p.iux,
                e.mag ? p.iax : -1);
  const double* vals[] = {frc->kk, frc->coef1, frc->coef2, frc->fda, frc->xyz0};
  for (const double* v : vals)
    for (int j = 0; j < 3; ++j)
      if (!std::isfinite(v[j])) return fail(PCMX_ERR_ARG, "pcmx_forcing: non-finite parameter");
  if (!std::isfinite(frc->phase) || !std::isfinite(frc->fact) || !std::isfinite(frc->k1_ff))
    return fail(PCMX_ERR_ARG, "pcmx_forcing: non-finite parameter");
  const int n[3] = {p.nx, p.ny, p.nz};
  const double d[3] = {p.dx, p.dy, p.dz};
  std::vector<double2> tab((size_t)p.nx + p.ny + p.nz + FORCE_NCOEF / 2);
  size_t o = 0;
  for (int a = 0; a < 3; ++a)
    for (int l = 0; l < n[a]; ++l) {
      // cell centres of the periodic grid (grid.f90:110-216); exp(cmplx(0., kk*k1_ff*x + phase))
      double c;
      {
#pragma clang fp contract(off)
        c = frc->xyz0[a] + (l + ((a == 2 && !p.lperi_z) ? 0.0 : 0.5)) * d[a];
      }
      const double th = frc->kk[a] * frc->k1_ff * c + (a == 0 ? frc->phase : 0.0);
      tab[o++] = make_double2(std::cos(th), std::sin(th));
    }
  double* cf = reinterpret_cast<double*>(tab.data() + o);
  for (int j = 0; j < 3; ++j) {
    cf[j] = frc->coef1[j];
    cf[3 + j] = frc->coef2[j];
    cf[6 + j] = frc->fda[j];
  }
  cf[9] = 0.0;
  if (!e.ftab) HIPCHECK(hipMalloc(&e.ftab, tab.size() * sizeof(double2)));
  HIPCHECK(hipMemcpyAsync(e.ftab, tab.data(), tab.size() * sizeof(double2), hipMemcpyHostToDevice, e.stream));
  HIPCHECK(hipStreamSynchronize(e.stream));
  e.frc = *frc;
  return PCMX_OK;
}

// ------------------------------------------------------------------ snapshots

int pcmx_write_record(const char* path, int append, const double* const* parts, const long* counts, int nparts,
                      long max_subrecord) {
  if (!path || nparts < 0 || (nparts > 0 && (!parts || !counts)) || max_subrecord < 0 ||
      max_subrecord > (long)FRecord::MAX_SUB)
    return fail(PCMX_ERR_ARG, "pcmx_write_record: bad arguments");
  uint64_t total = 0;
  for (int i = 0; i < nparts; ++i) {
    if (counts[i] < 0 || (counts[i] > 0 && !parts[i])) return fail(PCMX_ERR_ARG, "pcmx_write_record: part %d", i);
    total += (uint64_t)counts[i] * sizeof(double);
  }
  FILE* fp = std::fopen(path, append ? "ab" : "wb");
  if (!fp) return fail(PCMX_ERR_ARG, "cannot open %s: %s", path, std::strerror(errno));
  FRecord rec(fp, total, (uint64_t)max_subrecord);
  bool ok = true;
  for (int i = 0; i < nparts && ok; ++i) ok = rec.put(parts[i], (uint64_t)counts[i] * sizeof(double));
  ok = ok && rec.finish();
  ok = (std::fclose(fp) == 0) && ok;
  return ok ? PCMX_OK : fail(PCMX_ERR_ARG, "write to %s failed", path);
}

int pcmx_write_snapshot(const char* path, double t, const double* xyz0) {
  int rc = require();
  if (rc) return rc;
  Engine& e = *g;
  if (!path || !xyz0) return fail(PCMX_ERR_ARG, "pcmx_write_snapshot: path or xyz0 is NULL");
  const pcmx_params& p = e.par;
  const long mw = mw_of(p);
  constexpr long CHUNK = 8L << 20;  // doubles per pinned chunk (64 MiB)
  if (!e.staging) HIPCHECK(hipMalloc((void**)&e.staging, sizeof(double) * mw));
  for (int k = 0; k < 2; ++k) {
    if (!e.pin[k]) HIPCHECK(hipHostMalloc((void**)&e.pin[k], sizeof(double) * CHUNK, hipHostMallocDefault));
    if (!e.pin_ev[k]) HIPCHECK(hipEventCreateWithFlags(&e.pin_ev[k], hipEventDisableTiming));
  }
  std::unique_ptr<FILE, int (*)(FILE*)> fh(std::fopen(path, "wb"), &std::fclose);  // closed on every return
  if (!fh) return fail(PCMX_ERR_ARG, "cannot open %s: %s", path, std::strerror(errno));
  FILE* fp = fh.get();
  // record 1: f(mx,my,mz,1:mvar) with periodic ghosts (io_dist.f90:187)
  FRecord rec(fp, (uint64_t)mw * e.nvar * sizeof(double), 0);
  const double* state = current_state(e);
  if ((rc = apply_zbc(e, state, false))) return rc;  // z ghosts of the current state
  const long nxy = (long)p.nx * p.ny;
  const long nch = (mw + CHUNK - 1) / CHUNK;  // chunks per variable
  const long nall = nch * e.nvar;
  auto issue = [&](long c) -> int {  // pack (first chunk of a variable) and copy chunk c
    const int v = (int)(c / nch);
    const long off = (c % nch) * CHUNK, n = std::min(CHUNK, mw - off);
    if (c % nch == 0) {
      hipLaunchKernelGGL(k_pack_ghosts, dim3(1024), dim3(256), 0, e.stream, state + (long)v * e.nxyz, e.staging,
                         p.nx, p.ny, p.nz, e.zgh ? e.zgh + v * 6 * nxy : nullptr);
      HIPCHECK(hipGetLastError());
    }
    HIPCHECK(hipMemcpyAsync(e.pin[c & 1], e.staging + off, sizeof(double) * n, hipMemcpyDeviceToHost, e.stream));
    HIPCHECK(hipEventRecord(e.pin_ev[c & 1], e.stream));
    return PCMX_OK;
  };
  bool ok = true;
  if (nall > 0 && (rc = issue(0))) return rc;
  for (long c = 0; c < nall && ok; ++c) {
    // the copy of chunk c+1 runs while chunk c is written to the file; its
    // buffer held chunk c-1, already written
    if (c + 1 < nall && (rc = issue(c + 1))) return rc;
    HIPCHECK(hipEventSynchronize(e.pin_ev[c & 1]));
    const long n = std::min(CHUNK, mw - (c % nch) * CHUNK);
    ok = rec.put(e.pin[c & 1], (uint64_t)n * sizeof(double));
  }
  HIPCHECK(hipStreamSynchronize(e.stream));
  ok = ok && rec.finish();
  // record 2: t, x(mx), y(my), z(mz), dx, dy, dz (io_dist.f90:195).  The
  // periodic equidistant grid of grid.f90:110-216: x(i) = x0 + Lx*xi/nx with
  // xi = i - nghost - 1 + 1/2 (1-based i), i.e. x0 + (i - 3.5)*dx.
  const int mx = p.nx + 6, my = p.ny + 6, mz = p.nz + 6;
  std::vector<double> grid;
  grid.reserve(4 + mx + my + mz);
  grid.push_back(t);
  auto coord = [](double x0, int i, double d, double o) {
#pragma clang fp contract(off)
    return x0 + (i - o) * d;  // a rounded product, then the sum (no fused multiply-add)
  };
  // periodic: cell centres, x(l1) = x0 + dx/2; non-periodic z: z(n1) = z0
  for (int i = 1; i <= mx; ++i) grid.push_back(coord(xyz0[0], i, p.dx, 3.5));
  for (int i = 1; i <= my; ++i) grid.push_back(coord(xyz0[1], i, p.dy, 3.5));
  for (int i = 1; i <= mz; ++i) grid.push_back(coord(xyz0[2], i, p.dz, p.lperi_z ? 3.5 : 4.0));
  grid.push_back(p.dx);
  grid.push_back(p.dy);
  grid.push_back(p.dz);
  FRecord rec2(fp, grid.size() * sizeof(double), 0);
  ok = ok && rec2.put(grid.data(), grid.size() * sizeof(double)) && rec2.finish();
  ok = (std::fclose(fh.release()) == 0) && ok;
  return ok ? PCMX_OK : fail(PCMX_ERR_ARG, "write to %s failed", path);
}

}  // extern "C"
