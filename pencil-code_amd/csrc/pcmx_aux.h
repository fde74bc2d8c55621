// pcmx_aux.h -- the small kernels around the RHS kernels: the reduction
// finaliser, the explicit first-substep update for Courant-limited steps, the
// layout conversion between the reference f-array (with ghost zones) and
// device state, and the separate helical-forcing pass.  Included by the host
// translation unit only.
#pragma once

#include <hip/hip_runtime.h>

#include "pcmx_kernels.h"

namespace pcmx {

// Finalises the block partials in a fixed order (deterministic): urms, umax,
// brms (diagnostics.f90 sum_mn_name / max_mn_name with lsqrt, dVol_rel1 =
// 1/nwgrid at :274), dt1_max, and when set_dt the step dt = 1/dt1_max
// (sub.f90:4223-4254), and the box means.  One block of 256 threads.
__global__ void __launch_bounds__(256) k_finalize(const double* partials, long nblk, double* results,
                                                  double ncells, int mvar, int set_dt, int set_means) {
  __shared__ double lds[4][NRED];
  double r[NRED];
#pragma unroll
  for (int q = 0; q < NRED; ++q) r[q] = 0.0;
  for (long b = threadIdx.x; b < nblk; b += 256) {
#pragma unroll
    for (int q = 0; q < NRED; ++q) {
      const double v = partials[b * NRED + q];
      r[q] = (q == 1 || q == 3) ? fmax(r[q], v) : r[q] + v;
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NRED; ++q) {
    const double v = (q == 1 || q == 3) ? wave_max(r[q]) : wave_sum(r[q]);
    if (lane == 0) lds[wave][q] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double t[NRED];
    for (int q = 0; q < NRED; ++q) {
      t[q] = lds[0][q];
      for (int wv = 1; wv < 4; ++wv) t[q] = (q == 1 || q == 3) ? fmax(t[q], lds[wv][q]) : t[q] + lds[wv][q];
    }
    results[R_URMS] = sqrt(t[0] / ncells);
    results[R_UMAX] = sqrt(t[1]);
    results[R_BRMS] = sqrt(t[2] / ncells);
    results[R_DT1] = t[3];
    if (set_dt) results[R_DT] = 1.0 / t[3];
    if (set_means)
      for (int v = 0; v < mvar; ++v) results[R_MEAN0 + v] = t[4 + v] / ncells;
  }
}

// Deferred update of the first substep when dt comes from the Courant
// condition of that substep's RHS: f_out = f_in + (dt*beta)*w
// (timestep.f90:150,160-161).  Grid-stride over mvar*nxyz elements.
__global__ void __launch_bounds__(256) k_axpy_dt(const double* __restrict__ fin, double* __restrict__ fout,
                                                 const double* __restrict__ w, const double* d_dt, double beta,
                                                 long n) {
  const double dtb = d_dt[0] * beta;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256L) fout[i] = fin[i] + dtb * w[i];
}

// Device interior -> one variable of the reference f-array with periodic
// ghost zones (bc_per_x/y/z, boundcond.f90:2387-2433,2509-); with a
// non-periodic z, the z ghost planes come from gh (this variable's six
// ghost planes, k_zbc's layout).
__global__ void __launch_bounds__(256) k_pack_ghosts(const double* __restrict__ src, double* __restrict__ dst,
                                                     int nx, int ny, int nz, const double* __restrict__ gh) {
  const int mx = nx + 6, my = ny + 6, mz = nz + 6;
  const long mw = (long)mx * my * mz;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < mw; i += (long)gridDim.x * 256L) {
    const int ix = (int)(i % mx);
    const long r = i / mx;
    const int iy = (int)(r % my), iz = (int)(r / my);
    const int x = (ix - 3 + nx) % nx, y = (iy - 3 + ny) % ny;
    const int zi = iz - 3;
    if (gh && (zi < 0 || zi >= nz))
      dst[i] = gh[((long)(zi < 0 ? zi + 3 : zi - nz + 3) * ny + y) * nx + x];
    else
      dst[i] = src[((long)((zi + nz) % nz) * ny + y) * nx + x];
  }
}

// z boundary conditions of a non-periodic z (bc_sym_z, boundcond.f90:
// 3842-3900, codes of boundconds_z :1486-1503) on the device state f
// (interior, [v][z][y][x]): ghost planes -3..-1 and nz..nz+2 of every
// variable into gh ([v][6][ny][nx]) and, when zero_a, f = 0 on the boundary
// planes of 'a' variables, as boundconds_z does to the f-array before the
// RHS.  Bottom: 's' f(-i) = f(i), 'a' -f(i), 'a2' 2*f(0) - f(i); top (n2 =
// nz-1): 's' f(n2-i), 'a' -f(n2-i), 'a2' f(n2) + (f(n2) - f(n2-i)).
// The entropy slot may carry 'c1' or 'cT' (log density, no reference
// state), which read lnrho (slot 3) with its own rule applied first, as the
// variable loop of boundconds_z does:
//  'c1' (bc_ss_flux, entropy_bcs.f90:93-235, the K-profile branch):
//    cs2 = cs20*exp(gamma_m1*(lnrho(nb)-lnrho0)+cv1*ss(nb)), tmp = cp*FK/cs2,
//    bottom ss(-i) = ss(i) + (cp-cv)*((lnrho(i)-lnrho(-i)) + dz2_bound(-i)*tmp),
//    top ss(n2+i) = ss(n2-i) + (cp-cv)*(-(lnrho(n2+i)-lnrho(n2-i)) - dz2_bound(i)*tmp);
//  'cT' (bc_ss_temp_z, entropy_bcs.f90:1329-1457), tmpT = 2*cv*log(cs2b/cs20):
//    ss(nb) = 0.5*tmpT - (cp-cv)*(lnrho(nb)-lnrho0) (with zero_a),
//    ghost = -ss(mirror) + tmpT - (cp-cv)*(lnrho(mirror)+lnrho(ghost)-2*lnrho0).
struct ZBC {
  int bot[8], top[8];
  int iss;                            // entropy slot, or -1
  double cs20, gamma_m1, cv1, lnrho0, cp, cpmcv;
  double fk[2];                       // FbotKbot, FtopKtop
  double tmpT[2];                     // 2*cv*log(cs2bot/cs20), 2*cv*log(cs2top/cs20)
  double dz2[2][3];                   // dz2_bound(-i) (bottom), dz2_bound(i) (top), i = 1..3
};

__device__ __forceinline__ double sym_ghost(int c, double fb, double fi, bool top) {
  return c == PCMX_BC_S ? fi : (c == PCMX_BC_A ? -fi : (top ? fb + (fb - fi) : 2 * fb - fi));
}

__global__ void __launch_bounds__(256) k_zbc(double* __restrict__ f, double* __restrict__ gh, ZBC bc, int nvar,
                                             long nxy, int nz, int zero_a) {
  const long q = blockIdx.x * 256L + threadIdx.x;
  const int v = blockIdx.y;
  if (q >= nxy || v >= nvar) return;
  double* fv = f + (long)v * nxy * nz;
  double* gv = gh + (long)v * 6 * nxy;
  const int n2 = nz - 1;
  const double fb = fv[q], ft = fv[(long)n2 * nxy + q];
  const double* lr = f + 3L * nxy * nz;
#pragma unroll
  for (int side = 0; side < 2; ++side) {
    const int c = side ? bc.top[v] : bc.bot[v];
    const long nb = side ? n2 : 0;
    if (c <= PCMX_BC_A2) {
#pragma unroll
      for (int i = 1; i <= 3; ++i) {
        const double fi = fv[(nb + (side ? -i : i)) * nxy + q];
        gv[(long)(side ? 2 + i : 3 - i) * nxy + q] = sym_ghost(c, side ? ft : fb, fi, side != 0);
      }
      continue;
    }
    // entropy rules: lnrho's boundary value (0 under zero_a when its rule is
    // 'a', set before ss in the reference) and its ghosts
    const int lc = side ? bc.top[3] : bc.bot[3];
    const double lb = (zero_a && lc == PCMX_BC_A) ? 0.0 : lr[nb * nxy + q];
    const double sb = side ? ft : fb;
    if (c == PCMX_BC_C1) {
      const double cs2 = bc.cs20 * exp(bc.gamma_m1 * (lb - bc.lnrho0) + bc.cv1 * sb);
      const double tmp = bc.cp * bc.fk[side] / cs2;
#pragma unroll
      for (int i = 1; i <= 3; ++i) {
        const long in = (nb + (side ? -i : i)) * nxy + q;
        const double li = lr[in], lg = sym_ghost(lc, lb, li, side != 0);
        const double g = side ? fv[in] + bc.cpmcv * (-(lg - li) - bc.dz2[1][i - 1] * tmp)
                              : fv[in] + bc.cpmcv * ((li - lg) + bc.dz2[0][i - 1] * tmp);
        gv[(long)(side ? 2 + i : 3 - i) * nxy + q] = g;
      }
    } else {  // PCMX_BC_CT
      const double t = bc.tmpT[side];
      if (zero_a) fv[nb * nxy + q] = 0.5 * t - bc.cpmcv * (lb - bc.lnrho0);
#pragma unroll
      for (int i = 1; i <= 3; ++i) {
        const long in = (nb + (side ? -i : i)) * nxy + q;
        const double li = lr[in], lg = sym_ghost(lc, lb, li, side != 0);
        gv[(long)(side ? 2 + i : 3 - i) * nxy + q] = -fv[in] + t - bc.cpmcv * (li + lg - 2 * bc.lnrho0);
      }
    }
  }
  if (zero_a) {
    if (bc.bot[v] == PCMX_BC_A) fv[q] = 0.0;
    if (bc.top[v] == PCMX_BC_A) fv[(long)n2 * nxy + q] = 0.0;
  }
}

// One variable of the reference f-array -> device interior.
__global__ void __launch_bounds__(256) k_unpack_interior(const double* __restrict__ src, double* __restrict__ dst,
                                                         int nx, int ny, int nz) {
  const int mx = nx + 6, my = ny + 6;
  const long n = (long)nx * ny * nz;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256L) {
    const int x = (int)(i % nx);
    const long r = i / nx;
    const int y = (int)(r % ny), z = (int)(r / ny);
    dst[i] = src[((long)(z + 3) * my + (y + 3)) * mx + (x + 3)];
  }
}

// Helical forcing of one step as a separate pass over the forced vector
// (kernels without the fused epilogue); the term is force_term's.
__global__ void __launch_bounds__(256) k_add_forcing(double* __restrict__ f, const double2* __restrict__ tab,
                                                     double fact, int target, int nx, int ny, int nz) {
  const int l = blockIdx.x * 64 + (threadIdx.x & 63);
  const int m = blockIdx.y * 4 + (threadIdx.x >> 6);
  const int n = blockIdx.z;
  if (l >= nx || m >= ny) return;
  const long nxyz = (long)nx * ny * nz;
  double* p = f + target * nxyz + ((long)n * ny + m) * nx + l;
  double v[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) v[j] = p[j * nxyz];
  force_term(tab, fact, nx, ny, nz, l, m, n, v);
#pragma unroll
  for (int j = 0; j < 3; ++j) p[j * nxyz] = v[j];
}

}  // namespace pcmx
