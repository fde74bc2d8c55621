// pcmx_kernels.h -- what every RHS kernel variant shares: launch arguments,
// reduction slots and the wavefront / block reductions.
#pragma once

#include <hip/hip_runtime.h>

#include "pcmx_physics.h"

namespace pcmx {

// Reduction slots per block: 0 sum u2, 1 max u2, 2 sum b2, 3 max dt1,
// 4.. sum of each evolved variable (box means for the mixed form).
constexpr int NRED = 12;
constexpr int MAXVAR = 8;

// results[] written by k_finalize (device), read by the host on demand.
enum { R_URMS = 0, R_UMAX, R_BRMS, R_DT1, R_DT, R_MEAN0, R_COUNT = R_MEAN0 + MAXVAR };

enum { MODE_UPDATE = 0, MODE_W = 1, MODE_DT = 2 };

struct KArgs {
  const double* fin;    // state f at the start of the substep (interior, [v][z][y][x])
  double* fout;         // state after the substep
  double* w;            // 2N register (df), read when !first, written when write_w
  double* wout;         // where the split kernel stores the 2N register (w, or a
                        // second buffer in the deferred-update stage)
  const double* means;  // per-variable offsets of the mixed form (results + R_MEAN0)
  double* partials;     // per-block reduction slots (NRED each) or nullptr
  const double* d_dt;   // device dt (results + R_DT) for Courant-limited runs, else nullptr
  int nx, ny, nz, mvar, iss, iax;
  long nxyz;            // elements per variable
  double alpha;         // alpha_ts(itsub), timestep.f90:43
  double beta;          // beta_ts(itsub), timestep.f90:44
  double dtb;           // dt*beta_ts(itsub) when d_dt == nullptr
  int first;            // itsub == 1: df starts at zero (timestep.f90:110)
  int mode;             // MODE_*
  int write_w;          // store the 2N register (not on the last substep)
  // kernels with a deferred first-substep update (the RK stage comes in
  // `stage`): the second substep forms f1 = f0 + dt*beta1*w1 while staging
  int stage;
  double beta1;         // beta_ts(1)
  double dtb1;          // dt*beta_ts(1) when d_dt == nullptr
  // helical forcing added to the last substep's output (split kernel): the
  // force_term table or nullptr, fact of this step, first forced slot
  const double2* ftab;
  double ffact;
  int ftarget;
  // non-periodic z: ghost planes of fin ([var][6][ny][nx], planes -3..-1 then
  // nz..nz+2, filled by k_zbc before the launch), or nullptr for periodic z
  const double* zgh;
};

// Source of z plane p of variable v (p in -3 .. nz+2): the state, with
// periodic wrap, or the ghost-plane buffer of a non-periodic z.  Uniform
// per plane.
__device__ __forceinline__ const double* zsrc(const KArgs& a, int v, int p) {
  const long nxy = (long)a.nx * a.ny;
  if (a.zgh && (p < 0 || p >= a.nz)) return a.zgh + ((long)v * 6 + (p < 0 ? p + 3 : p - a.nz + 3)) * nxy;
  const int pz = p < 0 ? p + a.nz : (p >= a.nz ? p - a.nz : p);
  return a.fin + (long)v * a.nxyz + (long)pz * nxy;
}

// Helical forcing at cell (l, m, n) (forcing_hel, forcing.f90:2066-2474,
// default path: force_ampl = 1, cos(omega_ff t) = 1, force1_scl = 1).  tab
// holds the unit phasors exp(i(kk1 k1_ff x_l + phase)) (nx of them),
// exp(i kk2 k1_ff y_m) (ny) and exp(i kk3 k1_ff z_n) (nz) as (re, im),
// followed by coef1[3], coef2[3], fda[3] (FORCE_NCOEF doubles);
// fx = fact * phasor_x (fxyz_coefs_hel, :2003-2064), fxyz = (fx fy) fz, and
// component j of the force is fda_j Re((coef1_j + i coef2_j) fxyz).  f[j] is
// the state after the step; returns f[j] + force_j.  Written with explicit
// fused multiply-adds so that the separate forcing pass and the split
// kernel's fused epilogue round identically.
constexpr int FORCE_NCOEF = 10;  // 9 used, padded to whole double2 entries
__device__ __forceinline__ void force_term(const double2* tab, double fact, int nx, int ny, int nz, int l, int m,
                                           int n, double (&f)[3]) {
  const double2 px = tab[l], fy = tab[nx + m], fz = tab[nx + ny + n];
  const double* cf = reinterpret_cast<const double*>(tab + nx + ny + nz);
  const double fxr = px.x * fact, fxi = px.y * fact;
  const double ar = fma(fxr, fy.x, -(fxi * fy.y)), ai = fma(fxr, fy.y, fxi * fy.x);
  const double br = fma(ar, fz.x, -(ai * fz.y)), bi = fma(ar, fz.y, ai * fz.x);
#pragma unroll
  for (int j = 0; j < 3; ++j) f[j] = fma(cf[6 + j], fma(cf[j], br, -(cf[3 + j] * bi)), f[j]);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// Reduces the per-thread values over the block (wavefront shuffles, then
// LDS across waves) and writes NRED partials for this block.  Every thread
// of the block must call it.
template <int NTHREADS>
__device__ __forceinline__ void block_reduce_store(double (&r)[NRED], double* partials, long blk) {
  __shared__ double red_lds[NTHREADS / 64][NRED];
  const int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
  const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
  for (int q = 0; q < NRED; ++q) {
    const double v = (q == 1 || q == 3) ? wave_max(r[q]) : wave_sum(r[q]);
    if (lane == 0) red_lds[wave][q] = v;
  }
  __syncthreads();
  if (tid < NRED) {
    double v = red_lds[0][tid];
    for (int wv = 1; wv < NTHREADS / 64; ++wv)
      v = (tid == 1 || tid == 3) ? fmax(v, red_lds[wv][tid]) : v + red_lds[wv][tid];
    partials[blk * NRED + tid] = v;
  }
}

}  // namespace pcmx
