// pcmx_launch.h -- the interface between the host driver (pcmx_host.hip) and
// the kernel translation units (pcmx_kern_f64.hip, pcmx_kern_f32.hip), which
// hold the template instantiations of every RHS kernel variant and are
// compiled in parallel.
#pragma once

#include <hip/hip_runtime.h>

#include "pcmx_kernels.h"
#include "pcmx_zmarch.h"

namespace pcmx {

// Tile shapes of the split z-marching kernel.
enum Tile { TILE_32x8 = 0, TILE_16x16 = 1 };

// RK stages of the split kernel (see pcmx_zsplit.h).
enum Stage { ST_S1 = 0, ST_S2 = 1, ST_S3 = 2, ST_W = 3, ST_DT = 4, ST_S2D = 5 };

// Everything that selects a kernel variant for one launch.
struct LaunchCfg {
  int kernel;   // 1 direct, 2 z-marching, 3 split z-marching
  int tile;     // Tile (kernel 3)
  bool mag, ent;
  bool dt;      // kernels 1 and 2: evaluate the Courant terms
  ZPlan zp;     // kernels 2 and 3: the grid of tile columns
  int upw = 0;  // upwinding mask (UPW_UU, UPW_LNRHO, UPW_SS bits)
  bool lay = false;  // layer physics of samples/conv-slab (direct kernel only)
};

struct EnsBox;
struct EnsWork;

// Single box: launch the RHS kernel chosen by c for one RK stage (a.stage for
// kernel 3, a.mode for kernels 1 and 2).  Returns false for a variant this
// build does not provide.
bool launch_rhs_f64(const LaunchCfg& c, const KArgs& a, const Phys& P, hipStream_t s);
bool launch_rhs_f32(const LaunchCfg& c, const KArgs& a, const Phys& P, hipStream_t s);

// Ensemble: one persistent launch of nblocks blocks over the nwork work
// items, all boxes sharing the physics (mag, ent) of c and the RK stage.
bool launch_ens_f64(const LaunchCfg& c, int stage, const EnsBox* boxes, const EnsWork* work, int nwork,
                    double* partials, int nblocks, hipStream_t s);
bool launch_ens_f32(const LaunchCfg& c, int stage, const EnsBox* boxes, const EnsWork* work, int nwork,
                    double* partials, int nblocks, hipStream_t s);

// Tile of the split kernel for an nx x ny plane: 32 x 8 when it divides the
// plane, else 16 x 16, else -1 (no split kernel).
// Upwinding masks the split kernel provides on a tile (the others run on the
// direct kernel): uu + lnrho without entropy; lnrho + ss or uu + lnrho + ss
// with entropy on 32x8 tiles.
inline bool split_upw_ok(bool ent, int mask, int tile) {
  return mask == 0 || (!ent && mask == (UPW_UU | UPW_LNRHO)) ||
         (ent && tile == TILE_32x8 && (mask == (UPW_LNRHO | UPW_SS) || mask == (UPW_UU | UPW_LNRHO | UPW_SS)));
}

// The split kernel carries the layer physics for the samples/conv-slab
// choice: entropy, lnrho + ss upwinding, 32x8 tiles.
inline bool split_lay_ok(bool ent, int mask, int tile) {
  return ent && tile == TILE_32x8 && mask == (UPW_LNRHO | UPW_SS);
}

// Upwinding masks the direct kernel provides with the layer physics.
inline bool lay_upw_ok(bool ent, int mask) {
  if (ent) return mask == 0 || mask == UPW_LNRHO || mask == (UPW_LNRHO | UPW_SS) || mask == (UPW_UU | UPW_LNRHO | UPW_SS);
  return mask == 0 || mask == (UPW_UU | UPW_LNRHO);
}

inline int split_tile(int nx, int ny) {
  if ((long)nx * ny >= (1L << 31)) return -1;
  if (nx % 32 == 0 && ny % 8 == 0) return TILE_32x8;
  if (nx % 16 == 0 && ny % 16 == 0) return TILE_16x16;
  return -1;
}
inline int tile_tx(int tile) { return tile == TILE_16x16 ? 16 : 32; }
inline int tile_ty(int tile) { return tile == TILE_16x16 ? 16 : 8; }

// Blocks of the direct kernel for an nx x ny x nz grid.
inline long direct_blocks(int nx, int ny, int nz) { return (long)((nx + 63) / 64) * ((ny + 3) / 4) * nz; }

}  // namespace pcmx
