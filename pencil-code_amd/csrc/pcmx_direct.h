// pcmx_direct.h -- the direct fused RHS + RK kernel: one thread per cell,
// every stencil operand loaded from global memory with periodic wrap.
//
// It evaluates the stencils exactly as the reference writes them
// (deriv.f90 der_main :124-162, der2_main :497-550, derij_main bidiagonal
// :1400-1452, including the operand order of each bidiagonal bracket), so
// in FP64 it reproduces the reference arithmetic operation for operation
// apart from FMA contraction and the device exp.  It is the parity anchor
// for the z-marching kernel and the fallback for tiny or odd grids.
#pragma once

#include "pcmx_kernels.h"

namespace pcmx {

template <bool ENT, bool MAG>
struct Slots {
  static constexpr int NV = 4 + (ENT ? 1 : 0) + (MAG ? 3 : 0);
  static constexpr int ISS = ENT ? 4 : -1;
  static constexpr int IAX = MAG ? (ENT ? 5 : 4) : -1;
};

// LAY: the samples/conv-slab layer physics (Coriolis force, K-profile
// conduction, heating and cooling; Phys::zlay row of this z plane).
template <typename T, bool MAG, bool ENT, bool DT, int UPW = 0, bool LAY = false>
__global__ void __launch_bounds__(256) k_rhs_direct(KArgs a, Phys P) {
  using S = Slots<ENT, MAG>;
  constexpr int NV = S::NV;
  const int x = blockIdx.x * 64 + threadIdx.x;
  const int y = blockIdx.y * 4 + threadIdx.y;
  const int z = blockIdx.z;
  const bool active = (x < a.nx) && (y < a.ny);
  double r[NRED];
#pragma unroll
  for (int q = 0; q < NRED; ++q) r[q] = 0.0;

  if (active) {
    const long nxyz = a.nxyz;
    int xo[7];
    long yo[7], zo[7], zvs[7];
    const double* zb[7];  // plane z+k-3 of variable 0, and the variable stride there
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      xo[k] = (x + k - 3 + a.nx) % a.nx;
      yo[k] = (long)((y + k - 3 + a.ny) % a.ny) * a.nx;
      zo[k] = (long)((z + k - 3 + a.nz) % a.nz) * a.nx * a.ny;
      zb[k] = zsrc(a, 0, z + k - 3);
      zvs[k] = zsrc(a, 1, z + k - 3) - zb[k];
    }
    double mu[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) mu[v] = (sizeof(T) == 4) ? a.means[v] : 0.0;
    const double* fin = a.fin;
    // stencil operand: value at offset (i,j,k) in -3..3 of variable v, as an
    // offset from the box mean in the mixed form
    auto st = [&](int v, int i, int j, int k) -> T {
      const double val = zb[k + 3][v * zvs[k + 3] + yo[j + 3] + xo[i + 3]];
      if constexpr (sizeof(T) == 4) return (T)(val - mu[v]);
      else return val;
    };
    auto at = [&](int v, int dir, int o) -> T {
      return st(v, dir == 0 ? o : 0, dir == 1 ? o : 0, dir == 2 ? o : 0);
    };
    auto d1 = [&](int v, int dir) -> T {
      return (T)P.fac1[dir] * ((T)45.0 * (at(v, dir, 1) - at(v, dir, -1)) - (T)9.0 * (at(v, dir, 2) - at(v, dir, -2)) +
                               (at(v, dir, 3) - at(v, dir, -3)));
    };
    auto d2 = [&](int v, int dir) -> T {
      return (T)P.fac2[dir] * ((T)PCMX_D2C0 * at(v, dir, 0) + (T)PCMX_D2C1 * (at(v, dir, 1) + at(v, dir, -1)) +
                               (T)PCMX_D2C2 * (at(v, dir, 2) + at(v, dir, -2)) +
                               (T)PCMX_D2C3 * (at(v, dir, 3) + at(v, dir, -3)));
    };
    // bidiagonal mixed derivative, pair 0 = xy, 1 = xz, 2 = yz
    auto mix = [&](int v, int pair) -> T {
      T b[3];
#pragma unroll
      for (int q = 1; q <= 3; ++q) {
        if (pair == 0)       // f(x+q,y+q)-f(x-q,y+q)+f(x-q,y-q)-f(x+q,y-q)
          b[q - 1] = st(v, q, q, 0) - st(v, -q, q, 0) + st(v, -q, -q, 0) - st(v, q, -q, 0);
        else if (pair == 1)  // f(x+q,z+q)-f(x-q,z+q)+f(x-q,z-q)-f(x+q,z-q)
          b[q - 1] = st(v, q, 0, q) - st(v, -q, 0, q) + st(v, -q, 0, -q) - st(v, q, 0, -q);
        else                 // f(y+q,z+q)-f(y+q,z-q)+f(y-q,z-q)-f(y-q,z+q)
          b[q - 1] = st(v, 0, q, q) - st(v, 0, q, -q) + st(v, 0, -q, -q) - st(v, 0, -q, q);
      }
      const T fac = (T)(pair == 0 ? P.facxy : (pair == 1 ? P.facxz : P.facyz));
      return fac * ((T)270.0 * b[0] - (T)27.0 * b[1] + (T)2.0 * b[2]);
    };

    Derivs<T> d;
    const long c0 = zo[3] + yo[3] + xo[3];
    double fc[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) fc[v] = fin[v * nxyz + c0];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      d.u[i] = (T)fc[i];
#pragma unroll
      for (int j = 0; j < 3; ++j) d.uij[i][j] = d1(i, j);
      // D_xy u_i feeds graddiv for i = x, y; D_xz for i = x, z; D_yz for i = y, z
      const T mxy = (i != 2) ? mix(i, 0) : (T)0;
      const T mxz = (i != 1) ? mix(i, 1) : (T)0;
      const T myz = (i != 0) ? mix(i, 2) : (T)0;
      add_u_second(d, i, d2(i, 0), d2(i, 1), d2(i, 2), mxy, mxz, myz);
      d.glnrho[i] = d1(3, i);
    }
    d.lnrho = (T)fc[3];
    if constexpr (UPW != 0) {
      // upwind corrections: sum over j of |u_j| D6_j f, the sum in order x, y, z
      // (calc_del6_for_upwind, sub.f90:7508-7537)
      auto d6 = [&](int v, int dir) -> T {
        return der6_upwind<T>((T)P.fac1[dir], at(v, dir, 0), at(v, dir, 1) + at(v, dir, -1),
                              at(v, dir, 2) + at(v, dir, -2), at(v, dir, 3) + at(v, dir, -3));
      };
      const T ua[3] = {(T)fabs(fc[0]), (T)fabs(fc[1]), (T)fabs(fc[2])};
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        if (((v < 3) ? (UPW & UPW_UU) : (UPW & UPW_LNRHO)) == 0) continue;
        d.upw[v] = ua[0] * d6(v, 0) + ua[1] * d6(v, 1) + ua[2] * d6(v, 2);
      }
      if constexpr (ENT && (UPW & UPW_SS) != 0)
        d.upw[4] = ua[0] * d6(S::ISS, 0) + ua[1] * d6(S::ISS, 1) + ua[2] * d6(S::ISS, 2);
    }
    if constexpr (ENT) {
      d.ss = (T)fc[S::ISS];
#pragma unroll
      for (int j = 0; j < 3; ++j) d.gss[j] = d1(S::ISS, j);
      d.del2ss = d2(S::ISS, 0) + d2(S::ISS, 1) + d2(S::ISS, 2);
      d.del2lnrho = d2(3, 0) + d2(3, 1) + d2(3, 2);
    }
    if constexpr (MAG) {
      constexpr int IA = S::IAX;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        d.a[i] = (T)fc[IA + i];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          d.aij[i][j] = d1(IA + i, j);
          d.d2a[i][j] = d2(IA + i, j);
        }
      }
      d.mxa[0][0] = mix(IA + 0, 0);
      d.mxa[0][1] = mix(IA + 1, 0);
      d.mxa[1][0] = mix(IA + 0, 1);
      d.mxa[1][2] = mix(IA + 2, 1);
      d.mxa[2][1] = mix(IA + 1, 2);
      d.mxa[2][2] = mix(IA + 2, 2);
    }

    const long idx = ((long)z * a.ny + y) * a.nx + x;
    double w[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) w[v] = a.first ? 0.0 : a.alpha * a.w[v * nxyz + idx];
    double dt1 = 0.0, u2 = 0.0, b2 = 0.0;
    rhs_point<T, MAG, ENT, DT, UPW, LAY>(d, P, w, S::ISS, S::IAX, &dt1, &u2, &b2,
                                         LAY ? P.zlay + 4L * z : nullptr);

    if (a.mode == MODE_UPDATE) {
      const double dtb = a.d_dt ? a.d_dt[0] * a.beta : a.dtb;
#pragma unroll
      for (int v = 0; v < NV; ++v) a.fout[v * nxyz + idx] = fc[v] + dtb * w[v];
    }
    if (a.mode != MODE_DT && a.write_w) {
#pragma unroll
      for (int v = 0; v < NV; ++v) a.w[v * nxyz + idx] = w[v];
    }
    r[0] = u2;
    r[1] = u2;
    r[2] = b2;
    r[3] = dt1;
#pragma unroll
    for (int v = 0; v < NV; ++v) r[4 + v] = fc[v];
  }
  if (a.partials) {
    const long blk = blockIdx.x + (long)gridDim.x * (blockIdx.y + (long)gridDim.y * blockIdx.z);
    block_reduce_store<256>(r, a.partials, blk);
  }
}

}  // namespace pcmx
