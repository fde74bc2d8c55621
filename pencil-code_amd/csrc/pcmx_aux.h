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
struct ZBC {
  int bot[8], top[8];
};
__global__ void __launch_bounds__(256) k_zbc(double* __restrict__ f, double* __restrict__ gh, ZBC bc, int nvar,
                                             long nxy, int nz, int zero_a) {
  const long q = blockIdx.x * 256L + threadIdx.x;
  const int v = blockIdx.y;
  if (q >= nxy || v >= nvar) return;
  double* fv = f + (long)v * nxy * nz;
  double* gv = gh + (long)v * 6 * nxy;
  const int n2 = nz - 1;
  const double fb = fv[q], ft = fv[(long)n2 * nxy + q];
#pragma unroll
  for (int i = 1; i <= 3; ++i) {
    const double ib = fv[(long)i * nxy + q], it = fv[(long)(n2 - i) * nxy + q];
    const int cb = bc.bot[v], ct = bc.top[v];
    gv[(long)(3 - i) * nxy + q] = cb == PCMX_BC_S ? ib : (cb == PCMX_BC_A ? -ib : 2 * fb - ib);
    gv[(long)(2 + i) * nxy + q] = ct == PCMX_BC_S ? it : (ct == PCMX_BC_A ? -it : ft + (ft - it));
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
