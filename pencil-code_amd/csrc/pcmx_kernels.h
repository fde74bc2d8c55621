// pcmx_kernels.h -- device kernels shared by every RHS kernel variant:
// launch arguments, the wavefront reduction, the reduction finaliser, the
// deferred RK update for Courant-limited steps, and the layout conversion
// between the reference f-array (with ghost zones) and device state.
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
};

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
// ghost zones (bc_per_x/y/z, boundcond.f90:2387-2433,2509-).
__global__ void __launch_bounds__(256) k_pack_ghosts(const double* __restrict__ src, double* __restrict__ dst,
                                                     int nx, int ny, int nz) {
  const int mx = nx + 6, my = ny + 6, mz = nz + 6;
  const long mw = (long)mx * my * mz;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < mw; i += (long)gridDim.x * 256L) {
    const int ix = (int)(i % mx);
    const long r = i / mx;
    const int iy = (int)(r % my), iz = (int)(r / my);
    const int x = (ix - 3 + nx) % nx, y = (iy - 3 + ny) % ny, z = (iz - 3 + nz) % nz;
    dst[i] = src[((long)z * ny + y) * nx + x];
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

}  // namespace pcmx
