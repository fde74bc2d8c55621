// pcmx_zmarch.h -- the z-marching fused RHS + RK kernel (LDS plane ring).
//
// Each 256-thread block owns a 32 x 8 column of cells and marches it
// through a range of z planes.  Planes of every evolved variable, with a
// 3-cell x/y halo formed by periodic wrap, are staged once in an LDS ring of
// four planes (q..q+3 while plane q is produced); the next plane is
// prefetched into registers while the current one is computed.  x, y and
// xy-mixed derivatives read the output plane in LDS; z derivatives read the
// thread's own column from the ring plus three values kept in registers
// from earlier planes; the xz- and yz-mixed bidiagonal derivatives
// (deriv.f90:1400-1452) are split as
//   D_xz f(z) = fac * ( sum_k c_k g_k(z+k)  -  sum_k c_k g_k(z-k) ),
//   g_k(z') = f(x+k,z') - f(x-k,z'),   c = (270, -27, 2),
// where the first sum reads the ring and the second is accumulated in three
// registers while earlier planes pass through (the reference groups the
// same twelve terms as ((a-b)+c)-d per bracket; the results agree to
// rounding).  Each f value is read from HBM once per substep plus the halo
// rim shared with neighbouring tiles, which the L2 serves.
#pragma once

#include <cstdlib>

#include "../../include/pcmx.h"
#include "pcmx_direct.h"
#include "pcmx_kernels.h"

namespace pcmx {

constexpr int ZM_TX = 32, ZM_TY = 8, ZM_NT = ZM_TX * ZM_TY;
constexpr int ZM_PX = ZM_TX + 6, ZM_PY = ZM_TY + 6, ZM_PLANE = ZM_PX * ZM_PY;
constexpr int ZM_RING = 4;
constexpr int ZM_NPF = (ZM_PLANE + ZM_NT - 1) / ZM_NT;  // prefetch slots per thread

struct ZPlan {
  long nblocks = 0;
  int lz = 0, ntx = 0, nty = 0, nchunks = 0;
  int xcd_remap = 0;
};

inline bool zmarch_supported(const pcmx_params& p) {
  return p.nx % ZM_TX == 0 && p.ny % ZM_TY == 0 && p.nz >= 4 && (long)p.nx * p.ny < (1L << 31);
}

// Chooses the z-chunk length for tiles of tx x ty cells: the largest divisor
// of nz that still gives at least `target` blocks (8 per CU), but not below 8
// planes; lz_req > 0 (a divisor of nz) overrides.
inline ZPlan tile_plan(int nx, int ny, int nz, int tx, int ty, int lz_req) {
  ZPlan z;
  z.ntx = nx / tx;
  z.nty = ny / ty;
  const long ntiles = (long)z.ntx * z.nty;
  const long target = 2048;
  int lz = nz;
  if (lz_req > 0 && nz % lz_req == 0) {
    lz = lz_req;
  } else {
    int best = nz;
    for (int d = nz; d >= 1; --d) {
      if (nz % d) continue;
      if (d < 8 && d != nz) break;
      best = d;
      if (ntiles * (nz / d) >= target) break;
    }
    lz = best;
  }
  z.lz = lz;
  z.nchunks = nz / lz;
  z.nblocks = ntiles * z.nchunks;
  z.xcd_remap = (z.nblocks % 8 == 0) ? 1 : 0;
  return z;
}

// The first z-marching kernel's plan (32 x 8 tiles); PCMX_ZM_LZ overrides the
// chunk length (tuning sweeps).
inline ZPlan zmarch_plan(const pcmx_params& p) {
  const char* env = std::getenv("PCMX_ZM_LZ");
  return tile_plan(p.nx, p.ny, p.nz, ZM_TX, ZM_TY, env ? std::atoi(env) : 0);
}

template <bool MAG, bool ENT>
struct ZNeed {
  using S = Slots<ENT, MAG>;
  static constexpr int IA = S::IAX, ISS = S::ISS;
  __device__ static constexpr bool xy(int v) { return v == 0 || v == 1 || (MAG && (v == IA || v == IA + 1)); }
  __device__ static constexpr bool xz(int v) { return v == 0 || v == 2 || (MAG && (v == IA || v == IA + 2)); }
  __device__ static constexpr bool yz(int v) { return v == 1 || v == 2 || (MAG && (v == IA + 1 || v == IA + 2)); }
  __device__ static constexpr int ixz(int v) { return v == 0 ? 0 : (v == 2 ? 1 : (v == IA ? 2 : 3)); }
  __device__ static constexpr int iyz(int v) { return v == 1 ? 0 : (v == 2 ? 1 : (v == IA + 1 ? 2 : 3)); }
  __device__ static constexpr bool d2(int v) {
    return v < 3 || (MAG && v >= IA) || (ENT && (v == 3 || v == ISS));
  }
};

// Files the stencils of variable v at one cell into d (v follows the
// registration order uu, lnrho, [ss], [aa]; uu components must come in
// order x, y, z).
template <typename T, bool MAG, bool ENT>
__device__ __forceinline__ void file_derivs(Derivs<T>& d, int v, T dx, T dy, T dz, T dxx, T dyy, T dzz, T mxy,
                                            T mxz, T myz) {
  using S = Slots<ENT, MAG>;
  if (v < 3) {
    d.uij[v][0] = dx;
    d.uij[v][1] = dy;
    d.uij[v][2] = dz;
    add_u_second(d, v, dxx, dyy, dzz, mxy, mxz, myz);
  } else if (v == 3) {
    d.glnrho[0] = dx;
    d.glnrho[1] = dy;
    d.glnrho[2] = dz;
    d.del2lnrho = dxx + dyy + dzz;
  } else if (ENT && v == S::ISS) {
    d.gss[0] = dx;
    d.gss[1] = dy;
    d.gss[2] = dz;
    d.del2ss = dxx + dyy + dzz;
  } else if (MAG) {
    const int k = v - S::IAX;
    d.aij[k][0] = dx;
    d.aij[k][1] = dy;
    d.aij[k][2] = dz;
    d.d2a[k][0] = dxx;
    d.d2a[k][1] = dyy;
    d.d2a[k][2] = dzz;
    if (k != 2) d.mxa[0][k] = mxy;
    if (k != 1) d.mxa[1][k] = mxz;
    if (k != 0) d.mxa[2][k] = myz;
  }
}

template <typename T, bool MAG, bool ENT, bool DT>
__global__ void __launch_bounds__(ZM_NT, 1) k_rhs_zmarch(KArgs a, Phys P, ZPlan zp) {
  using S = Slots<ENT, MAG>;
  using N = ZNeed<MAG, ENT>;
  constexpr int NV = S::NV, IA = S::IAX, ISS = S::ISS;
  constexpr int SLOT = NV * ZM_PLANE;
  constexpr int NXZ = MAG ? 4 : 2;
  __shared__ T lds[ZM_RING * SLOT];

  const int tid = threadIdx.x;
  const int tx = tid % ZM_TX, ty = tid / ZM_TX;
  // XCD-aware work order: consecutive work items (neighbouring tiles that
  // share halo rims) go to the same XCD, whose L2 then serves the rims.
  const long b = blockIdx.x;
  const long wid = zp.xcd_remap ? (b % 8) * (zp.nblocks / 8) + b / 8 : b;
  const int txt = (int)(wid % zp.ntx);
  const long rest = wid / zp.ntx;
  const int tyt = (int)(rest % zp.nty);
  const int chunk = (int)(rest / zp.nty);
  const int x0 = txt * ZM_TX, y0 = tyt * ZM_TY, z0 = chunk * zp.lz;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const long nxy = (long)nx * ny, nxyz = a.nxyz;

  // plane staging geometry (identical for every plane and variable)
  int goff[ZM_NPF], loff[ZM_NPF];
  bool pvalid[ZM_NPF];
#pragma unroll
  for (int s = 0; s < ZM_NPF; ++s) {
    const int e = tid + s * ZM_NT;
    pvalid[s] = e < ZM_PLANE;
    const int ee = pvalid[s] ? e : 0;
    const int row = ee / ZM_PX, col = ee % ZM_PX;
    int gx = x0 - 3 + col, gy = y0 - 3 + row;
    gx += (gx < 0) ? nx : 0;
    gx -= (gx >= nx) ? nx : 0;
    gy += (gy < 0) ? ny : 0;
    gy -= (gy >= ny) ? ny : 0;
    goff[s] = gy * nx + gx;
    loff[s] = ee;
  }
  double mu[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) mu[v] = (sizeof(T) == 4) ? a.means[v] : 0.0;

  auto zoff = [&](int p) -> long {
    int zz = p % nz;
    zz += (zz < 0) ? nz : 0;
    return (long)zz * nxy;
  };
  auto slot_of = [&](int p) -> int { return (p - z0 + 3) & 3; };

  double pre[NV][ZM_NPF];
  auto load_plane = [&](int p) {
    const double* src = a.fin + zoff(p);
#pragma unroll
    for (int v = 0; v < NV; ++v)
#pragma unroll
      for (int s = 0; s < ZM_NPF; ++s)
        if (pvalid[s]) pre[v][s] = src[v * nxyz + goff[s]];
  };
  auto store_plane = [&](int p) {
    T* dst = lds + slot_of(p) * SLOT;
#pragma unroll
    for (int v = 0; v < NV; ++v)
#pragma unroll
      for (int s = 0; s < ZM_NPF; ++s)
        if (pvalid[s]) {
          if constexpr (sizeof(T) == 4) dst[v * ZM_PLANE + loff[s]] = (T)(pre[v][s] - mu[v]);
          else dst[v * ZM_PLANE + loff[s]] = pre[v][s];
        }
  };

  // prime the ring with planes z0-3 .. z0
  for (int p = z0 - 3; p <= z0; ++p) {
    load_plane(p);
    store_plane(p);
  }
  __syncthreads();

  T back[NV][3];
  T mxz[NXZ][3], myz[NXZ][3];
#pragma unroll
  for (int v = 0; v < NV; ++v) back[v][0] = back[v][1] = back[v][2] = (T)0;
#pragma unroll
  for (int i = 0; i < NXZ; ++i) {
    mxz[i][0] = mxz[i][1] = mxz[i][2] = (T)0;
    myz[i][0] = myz[i][1] = myz[i][2] = (T)0;
  }
  double red[NRED];
#pragma unroll
  for (int q = 0; q < NRED; ++q) red[q] = 0.0;

  const int c = (ty + 3) * ZM_PX + (tx + 3);
  const int x = x0 + tx, y = y0 + ty;
  const long own = (long)y * nx + x;
  const int zlast = z0 + zp.lz - 1;
  const T fx1 = (T)P.fac1[0], fy1 = (T)P.fac1[1], fz1 = (T)P.fac1[2];
  const T fx2 = (T)P.fac2[0], fy2 = (T)P.fac2[1], fz2 = (T)P.fac2[2];
  const T C0 = (T)PCMX_D2C0, C1 = (T)PCMX_D2C1, C2 = (T)PCMX_D2C2, C3 = (T)PCMX_D2C3;

  for (int q = z0 - 3; q <= zlast; ++q) {
    const bool emit = q >= z0;
    const bool fetch = q + 4 <= zlast + 3;
    if (fetch) load_plane(q + 4);
    double wv[NV], fown[NV];
    const long zo = (long)q * nxy;  // output planes never wrap
    if (emit) {
#pragma unroll
      for (int v = 0; v < NV; ++v) wv[v] = a.first ? 0.0 : a.alpha * a.w[v * nxyz + zo + own];
      if constexpr (sizeof(T) == 4) {
#pragma unroll
        for (int v = 0; v < NV; ++v) fown[v] = a.fin[v * nxyz + zo + own];
      }
    }
    const T* L0 = lds + slot_of(q) * SLOT + c;
    const T* L1 = lds + slot_of(q + 1) * SLOT + c;
    const T* L2 = lds + slot_of(q + 2) * SLOT + c;
    const T* L3 = lds + slot_of(q + 3) * SLOT + c;
    Derivs<T> d;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      const T* p0 = L0 + v * ZM_PLANE;
      const T c0 = p0[0];
      if constexpr (sizeof(T) == 8) fown[v] = (double)c0;
      if (emit) {
        const T* p1 = L1 + v * ZM_PLANE;
        const T* p2 = L2 + v * ZM_PLANE;
        const T* p3 = L3 + v * ZM_PLANE;
        const T xp1 = p0[1], xm1 = p0[-1], xp2 = p0[2], xm2 = p0[-2], xp3 = p0[3], xm3 = p0[-3];
        const T yp1 = p0[ZM_PX], ym1 = p0[-ZM_PX], yp2 = p0[2 * ZM_PX], ym2 = p0[-2 * ZM_PX];
        const T yp3 = p0[3 * ZM_PX], ym3 = p0[-3 * ZM_PX];
        const T zp1 = p1[0], zp2 = p2[0], zp3 = p3[0];
        const T zm1 = back[v][0], zm2 = back[v][1], zm3 = back[v][2];
        // der_main (deriv.f90:124-162)
        const T dx = fx1 * ((T)45.0 * (xp1 - xm1) - (T)9.0 * (xp2 - xm2) + (xp3 - xm3));
        const T dy = fy1 * ((T)45.0 * (yp1 - ym1) - (T)9.0 * (yp2 - ym2) + (yp3 - ym3));
        const T dz = fz1 * ((T)45.0 * (zp1 - zm1) - (T)9.0 * (zp2 - zm2) + (zp3 - zm3));
        T dxx = 0, dyy = 0, dzz = 0;
        if (N::d2(v)) {  // der2_main (deriv.f90:497-550)
          dxx = fx2 * (C0 * c0 + C1 * (xp1 + xm1) + C2 * (xp2 + xm2) + C3 * (xp3 + xm3));
          dyy = fy2 * (C0 * c0 + C1 * (yp1 + ym1) + C2 * (yp2 + ym2) + C3 * (yp3 + ym3));
          dzz = fz2 * (C0 * c0 + C1 * (zp1 + zm1) + C2 * (zp2 + zm2) + C3 * (zp3 + zm3));
        }
        T mxy = 0, mxzv = 0, myzv = 0;
        if (N::xy(v)) {  // bidiagonal xy, operand order of deriv.f90:1410-1418
          const T b1 = p0[1 + ZM_PX] - p0[-1 + ZM_PX] + p0[-1 - ZM_PX] - p0[1 - ZM_PX];
          const T b2 = p0[2 + 2 * ZM_PX] - p0[-2 + 2 * ZM_PX] + p0[-2 - 2 * ZM_PX] - p0[2 - 2 * ZM_PX];
          const T b3 = p0[3 + 3 * ZM_PX] - p0[-3 + 3 * ZM_PX] + p0[-3 - 3 * ZM_PX] - p0[3 - 3 * ZM_PX];
          mxy = (T)P.facxy * ((T)270.0 * b1 - (T)27.0 * b2 + (T)2.0 * b3);
        }
        if (N::xz(v)) {
          const T plus = (T)270.0 * (p1[1] - p1[-1]) - (T)27.0 * (p2[2] - p2[-2]) + (T)2.0 * (p3[3] - p3[-3]);
          mxzv = (T)P.facxz * (plus - mxz[N::ixz(v)][0]);
        }
        if (N::yz(v)) {
          const T plus = (T)270.0 * (p1[ZM_PX] - p1[-ZM_PX]) - (T)27.0 * (p2[2 * ZM_PX] - p2[-2 * ZM_PX]) +
                         (T)2.0 * (p3[3 * ZM_PX] - p3[-3 * ZM_PX]);
          myzv = (T)P.facyz * (plus - myz[N::iyz(v)][0]);
        }
        const T val = (sizeof(T) == 4) ? (T)fown[v] : c0;
        if (v < 3) d.u[v] = val;
        else if (v == 3) d.lnrho = val;
        else if (ENT && v == ISS) d.ss = val;
        else if (MAG) d.a[v - IA] = val;
        file_derivs<T, MAG, ENT>(d, v, dx, dy, dz, dxx, dyy, dzz, mxy, mxzv, myzv);
      }
      // carry plane q forward: minus terms of the z-mixed derivatives of the
      // next three planes, and the z-window of the own column
      if (N::xz(v)) {
        T* m = mxz[N::ixz(v)];
        const T g1 = p0[1] - p0[-1], g2 = p0[2] - p0[-2], g3 = p0[3] - p0[-3];
        m[0] = m[1] + (T)270.0 * g1;
        m[1] = m[2] - (T)27.0 * g2;
        m[2] = (T)2.0 * g3;
      }
      if (N::yz(v)) {
        T* m = myz[N::iyz(v)];
        const T h1 = p0[ZM_PX] - p0[-ZM_PX], h2 = p0[2 * ZM_PX] - p0[-2 * ZM_PX];
        const T h3 = p0[3 * ZM_PX] - p0[-3 * ZM_PX];
        m[0] = m[1] + (T)270.0 * h1;
        m[1] = m[2] - (T)27.0 * h2;
        m[2] = (T)2.0 * h3;
      }
      back[v][2] = back[v][1];
      back[v][1] = back[v][0];
      back[v][0] = c0;
    }
    if (emit) {
      double dt1 = 0.0, u2 = 0.0, b2 = 0.0;
      rhs_point<T, MAG, ENT, DT>(d, P, wv, ISS, IA, &dt1, &u2, &b2);
      if (a.mode == MODE_UPDATE) {
        const double dtb = a.d_dt ? a.d_dt[0] * a.beta : a.dtb;
#pragma unroll
        for (int v = 0; v < NV; ++v) a.fout[v * nxyz + zo + own] = fown[v] + dtb * wv[v];
      }
      if (a.mode != MODE_DT && a.write_w) {
#pragma unroll
        for (int v = 0; v < NV; ++v) a.w[v * nxyz + zo + own] = wv[v];
      }
      red[0] += u2;
      red[1] = fmax(red[1], u2);
      red[2] += b2;
      red[3] = fmax(red[3], dt1);
#pragma unroll
      for (int v = 0; v < NV; ++v) red[4 + v] += fown[v];
    }
    __syncthreads();
    if (fetch) store_plane(q + 4);
    __syncthreads();
  }
  if (a.partials) block_reduce_store<ZM_NT>(red, a.partials, (long)blockIdx.x);
}

template <typename T, bool MAG, bool ENT, bool DT>
void launch_zmarch(const KArgs& a, const Phys& P, const ZPlan& zp, hipStream_t s) {
  hipLaunchKernelGGL((k_rhs_zmarch<T, MAG, ENT, DT>), dim3((unsigned)zp.nblocks), dim3(ZM_NT), 0, s, a, P, zp);
}

}  // namespace pcmx
