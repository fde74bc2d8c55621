// pcmx_dispatch.h -- template dispatch from a LaunchCfg to the kernel
// variants; included by the kernel translation units (pcmx_kern_phys.hip
// instantiates launch_phys for one precision and physics, pcmx_kern_f64.hip
// and pcmx_kern_f32.hip the per-precision entry points and the ensemble).
#pragma once

#include "pcmx_direct.h"
#include "pcmx_launch.h"
#include "pcmx_zmarch.h"
#include "pcmx_zsplit.h"

namespace pcmx {

template <typename T, bool MAG, bool ENT, bool DT, int UPW, bool LAY = false>
void launch_direct(const KArgs& a, const Phys& P, hipStream_t s) {
  dim3 blk(64, 4, 1), grd((a.nx + 63) / 64, (a.ny + 3) / 4, a.nz);
  hipLaunchKernelGGL((k_rhs_direct<T, MAG, ENT, DT, UPW, LAY>), grd, blk, 0, s, a, P);
}

template <typename T, bool MAG, bool ENT, class TL, int UPW, bool LAY = false>
bool launch_split(const KArgs& a, const Phys& P, const ZPlan& zp, hipStream_t s) {
  switch (a.stage) {
    case ST_S1: launch_zsplit<T, MAG, ENT, ST_S1, TL, UPW, LAY>(a, P, zp, s); return true;
    case ST_S2: launch_zsplit<T, MAG, ENT, ST_S2, TL, UPW, LAY>(a, P, zp, s); return true;
    case ST_S3: launch_zsplit<T, MAG, ENT, ST_S3, TL, UPW, LAY>(a, P, zp, s); return true;
    case ST_W: launch_zsplit<T, MAG, ENT, ST_W, TL, UPW, LAY>(a, P, zp, s); return true;
    case ST_S2D: launch_zsplit<T, MAG, ENT, ST_S2D, TL, UPW, LAY>(a, P, zp, s); return true;
    case ST_DT: launch_zsplit<T, MAG, ENT, ST_DT, TL, UPW, LAY>(a, P, zp, s); return true;
    default: return false;
  }
}

template <typename T, bool MAG, bool ENT, int UPW, bool LAY = false>
void launch_direct_dt(const LaunchCfg& c, const KArgs& a, const Phys& P, hipStream_t s) {
  if (c.dt) launch_direct<T, MAG, ENT, true, UPW, LAY>(a, P, s);
  else launch_direct<T, MAG, ENT, false, UPW, LAY>(a, P, s);
}

// The layer physics (c.lay) runs on the direct kernel with the upwinding
// masks of samples/conv-slab and its neighbours: none, lnrho, lnrho + ss,
// uu + lnrho + ss with entropy; none or uu + lnrho without (lay_upw_ok).
template <typename T, bool MAG, bool ENT>
bool launch_direct_lay(const LaunchCfg& c, const KArgs& a, const Phys& P, hipStream_t s) {
  switch (c.upw) {
    case 0: launch_direct_dt<T, MAG, ENT, 0, true>(c, a, P, s); return true;
    default: break;
  }
  if constexpr (ENT) {
    switch (c.upw) {
      case UPW_LNRHO: launch_direct_dt<T, MAG, ENT, UPW_LNRHO, true>(c, a, P, s); return true;
      case UPW_LNRHO | UPW_SS: launch_direct_dt<T, MAG, ENT, UPW_LNRHO | UPW_SS, true>(c, a, P, s); return true;
      case UPW_UU | UPW_LNRHO | UPW_SS:
        launch_direct_dt<T, MAG, ENT, UPW_UU | UPW_LNRHO | UPW_SS, true>(c, a, P, s);
        return true;
      default: return false;
    }
  } else {
    if (c.upw == (UPW_UU | UPW_LNRHO)) {
      launch_direct_dt<T, MAG, ENT, UPW_UU | UPW_LNRHO, true>(c, a, P, s);
      return true;
    }
    return false;
  }
}

// Upwinding: every mask in the direct kernel; in the split kernel uu and
// lnrho together without entropy (the published configuration), and with
// entropy lnrho + ss or uu + lnrho + ss (split_upw_ok); none in the first
// z-marching kernel.
template <typename T, bool MAG, bool ENT>
bool launch_phys(const LaunchCfg& c, const KArgs& a, const Phys& P, hipStream_t s) {
  if (c.lay) {
    if (c.kernel == 3) {  // entropy, lnrho + ss upwinding, 32x8 tiles (split_lay_ok)
      if constexpr (ENT) {
        constexpr int M6 = UPW_LNRHO | UPW_SS;
        if (c.tile == TILE_32x8 && c.upw == M6) return launch_split<T, MAG, ENT, ZT32x8, M6, true>(a, P, c.zp, s);
      }
      return false;
    }
    return c.kernel == 1 && launch_direct_lay<T, MAG, ENT>(c, a, P, s);
  }
  if (c.kernel == 3) {
    if (c.upw == 0) {
      if (c.tile == TILE_16x16) return launch_split<T, MAG, ENT, ZT16x16, 0>(a, P, c.zp, s);
      return launch_split<T, MAG, ENT, ZT32x8, 0>(a, P, c.zp, s);
    }
    if constexpr (!ENT) {
      if (c.upw != (UPW_UU | UPW_LNRHO)) return false;
      if (c.tile == TILE_16x16) return launch_split<T, MAG, ENT, ZT16x16, UPW_UU | UPW_LNRHO>(a, P, c.zp, s);
      return launch_split<T, MAG, ENT, ZT32x8, UPW_UU | UPW_LNRHO>(a, P, c.zp, s);
    } else {
      // 32x8 tiles only (split_upw_ok)
      constexpr int M6 = UPW_LNRHO | UPW_SS, M7 = UPW_UU | UPW_LNRHO | UPW_SS;
      if (c.tile != TILE_32x8) return false;
      if (c.upw == M6) return launch_split<T, MAG, ENT, ZT32x8, M6>(a, P, c.zp, s);
      if (c.upw == M7) return launch_split<T, MAG, ENT, ZT32x8, M7>(a, P, c.zp, s);
      return false;
    }
  }
  if (c.kernel == 2) {
    if (c.upw != 0) return false;
    if (c.dt) launch_zmarch<T, MAG, ENT, true>(a, P, c.zp, s);
    else launch_zmarch<T, MAG, ENT, false>(a, P, c.zp, s);
    return true;
  }
  switch (c.upw) {
    case 0: launch_direct_dt<T, MAG, ENT, 0>(c, a, P, s); return true;
    case UPW_UU: launch_direct_dt<T, MAG, ENT, UPW_UU>(c, a, P, s); return true;
    case UPW_LNRHO: launch_direct_dt<T, MAG, ENT, UPW_LNRHO>(c, a, P, s); return true;
    case UPW_UU | UPW_LNRHO: launch_direct_dt<T, MAG, ENT, UPW_UU | UPW_LNRHO>(c, a, P, s); return true;
    default: break;
  }
  if constexpr (ENT) {  // lupw_ss: entropy runs only
    switch (c.upw) {
      case UPW_SS: launch_direct_dt<T, MAG, ENT, UPW_SS>(c, a, P, s); return true;
      case UPW_UU | UPW_SS: launch_direct_dt<T, MAG, ENT, UPW_UU | UPW_SS>(c, a, P, s); return true;
      case UPW_LNRHO | UPW_SS: launch_direct_dt<T, MAG, ENT, UPW_LNRHO | UPW_SS>(c, a, P, s); return true;
      case UPW_UU | UPW_LNRHO | UPW_SS:
        launch_direct_dt<T, MAG, ENT, UPW_UU | UPW_LNRHO | UPW_SS>(c, a, P, s);
        return true;
      default: break;
    }
  }
  return false;
}

// The instantiations of launch_phys live in pcmx_kern_phys.hip, one unit per
// (precision, MAG, ENT); the units that only dispatch declare them extern.
#define PCMX_EXTERN_PHYS(T)                                                                 \
  extern template bool launch_phys<T, false, false>(const LaunchCfg&, const KArgs&, const Phys&, hipStream_t); \
  extern template bool launch_phys<T, false, true>(const LaunchCfg&, const KArgs&, const Phys&, hipStream_t);  \
  extern template bool launch_phys<T, true, false>(const LaunchCfg&, const KArgs&, const Phys&, hipStream_t);  \
  extern template bool launch_phys<T, true, true>(const LaunchCfg&, const KArgs&, const Phys&, hipStream_t);

template <typename T>
bool launch_rhs_t(const LaunchCfg& c, const KArgs& a, const Phys& P, hipStream_t s) {
  if (c.mag) return c.ent ? launch_phys<T, true, true>(c, a, P, s) : launch_phys<T, true, false>(c, a, P, s);
  return c.ent ? launch_phys<T, false, true>(c, a, P, s) : launch_phys<T, false, false>(c, a, P, s);
}

// ---------------------------------------------------------------- ensemble
template <typename T, bool MAG, bool ENT, bool POLY, int ST, class TL>
void launch_ens_k(const EnsBox* boxes, const EnsWork* work, int nwork, double* partials, int nblocks,
                  hipStream_t s) {
  hipLaunchKernelGGL((k_rhs_zsplit_ens<T, MAG, ENT, POLY, ST, TL>), dim3((unsigned)nblocks),
                     dim3(zs_threads<MAG>()), 0, s, boxes, work, nwork, partials);
}

template <typename T, bool MAG, bool ENT, bool POLY, class TL>
bool launch_ens_st(int stage, const EnsBox* b, const EnsWork* w, int nw, double* pr, int nb, hipStream_t s) {
  switch (stage) {
    case ST_S1: launch_ens_k<T, MAG, ENT, POLY, ST_S1, TL>(b, w, nw, pr, nb, s); return true;
    case ST_S2: launch_ens_k<T, MAG, ENT, POLY, ST_S2, TL>(b, w, nw, pr, nb, s); return true;
    case ST_S3: launch_ens_k<T, MAG, ENT, POLY, ST_S3, TL>(b, w, nw, pr, nb, s); return true;
    case ST_DT: launch_ens_k<T, MAG, ENT, POLY, ST_DT, TL>(b, w, nw, pr, nb, s); return true;
    default: return false;  // Courant-limited ensembles are not provided
  }
}

// The ensemble provides isothermal hydro and MHD (fixed steps).
template <typename T>
bool launch_ens_t(const LaunchCfg& c, int stage, const EnsBox* b, const EnsWork* w, int nw, double* pr, int nb,
                  hipStream_t s) {
  if (c.ent) return false;
  if (c.tile == TILE_16x16)
    return c.mag ? launch_ens_st<T, true, false, false, ZT16x16>(stage, b, w, nw, pr, nb, s)
                 : launch_ens_st<T, false, false, false, ZT16x16>(stage, b, w, nw, pr, nb, s);
  return c.mag ? launch_ens_st<T, true, false, false, ZT32x8>(stage, b, w, nw, pr, nb, s)
               : launch_ens_st<T, false, false, false, ZT32x8>(stage, b, w, nw, pr, nb, s);
}

}  // namespace pcmx
