// pcmx_zsplit.h -- z-marching fused RHS + RK kernel with the flow and field
// halves of the RHS on separate wavefronts.
//
// Tiling and stencil scheme are those of the first z-marching kernel: a block
// owns a 32 x 8 column of cells and marches it through a range of z planes;
// planes of every evolved variable with a 3-cell periodic x/y halo sit in an
// LDS ring of four (planes q..q+3 while q is produced); z derivatives use the
// thread's own column (three planes ahead from the ring, three behind in
// registers); the xz and yz bidiagonal mixed derivatives (deriv.f90:1400-1452)
// are split as
//   D_xz f(z) = fac * ( sum_k c_k g_k(z+k) - sum_k c_k g_k(z-k) ),
//   g_k(z') = f(x+k,z') - f(x-k,z'),  c = (270, -27, 2),
// the first sum read from the ring, the second carried in registers.
//
// What differs:
//  * With a magnetic field the block has 512 threads: waves 0-3 evaluate the
//    flow half of the RHS (uu, lnrho, ss: hydro, density, energy, viscosity)
//    and waves 4-7 the field half (aa: B, J, J x B / rho, induction) for the
//    same 256 cells.  The field half passes J x B / rho, J^2 and the Alfven
//    Courant term through LDS.  One block per CU then runs 8 waves (2 per
//    SIMD) instead of 4, and each thread holds the registers of half the
//    variables.
//  * The RK stage is a template parameter, so every memory operation of the
//    main loop is unconditional and the compiler can count outstanding loads
//    exactly.
//  * Software pipeline, one plane deep for every load: while plane q is
//    computed, plane q+4 (all variables, with halo) and the 2N register of
//    the own cells of plane q+1 are in flight; each is consumed one phase
//    after it was issued.
//  * Data touched once (the own-cell 2N register loads, every state and 2N
//    register store) use non-temporal accesses, so that the L2 keeps the
//    halo rims that neighbouring tiles read again.
#pragma once

#include <type_traits>

#include "pcmx_direct.h"
#include "pcmx_kernels.h"
#include "pcmx_launch.h"
#include "pcmx_zmarch.h"

namespace pcmx {

// RK stages of the 2N scheme (timestep.f90:110-161) as the fused kernel sees
// them.  S1: first substep with a fixed step (store f1 = f0 + dt*beta1*w1 and
// w1); S2: middle substep (reads own-cell w, stores f and w); S3: last
// substep (reads w, stores f); DT: reductions only.
// Courant-limited steps defer the first update: W stores only w1 = RHS(f0)
// and reduces dt1 (alpha_1 = 0), and S2D stages f1 = f0 + dt*beta1*w1 into
// the LDS ring as it loads f0 and w1 (the operation of timestep.f90:160, so
// f1 is bitwise the value S1 would store) and writes w2 to a second buffer
// (neighbouring blocks still read w1 in their halos).  That replaces the
// separate update pass an explicit Courant step needs once dt is known:
// 9 state accesses per variable and step plus the w1 halo, against 12.  With
// a fixed step the explicit form is cheaper (w1 would be read with halo).
template <int ST>
struct StageT {
  static constexpr bool FORM1 = ST == ST_S2D;
  static constexpr bool READW = ST == ST_S2 || ST == ST_S2D || ST == ST_S3;
  static constexpr bool WRITEW = ST == ST_S1 || ST == ST_S2 || ST == ST_S2D || ST == ST_W;
  static constexpr bool UPDATE = ST == ST_S1 || ST == ST_S2 || ST == ST_S2D || ST == ST_S3;
  static constexpr bool DT = ST == ST_W || ST == ST_DT;
  static constexpr bool DIAG = ST == ST_S1 || ST == ST_W || ST == ST_DT;
};

// Tile geometry: a wave group of 256 threads owns TX x TY cells of a z
// plane; the staged plane carries a 3-cell halo on each side.
template <int TX_, int TY_>
struct ZTile {
  static constexpr int TX = TX_, TY = TY_, NT = TX_ * TY_, PX = TX_ + 6, PY = TY_ + 6, PLANE = PX * PY;
  static_assert(NT == 256, "one thread per cell of the tile, 256 per wave group");
};
using ZT32x8 = ZTile<32, 8>;
using ZT16x16 = ZTile<16, 16>;

template <bool MAG>
constexpr int zs_threads() {
  return MAG ? 512 : 256;
}

// Cells per staged plane in LDS (VS), and the staging slots: K slots per
// staged plane at cells tid + k*NTOT (a wave-uniform base plus a 32-bit lane
// offset) and, when the remainders fit (PACK), one more pass for them: with
// pair planes, wave 0 stages the remainders of the NP pair planes and wave 1
// those of the NS single planes (so the pair/single choice is uniform per
// wave); with single planes only, threads tid < NS*R.  Otherwise the slots
// cover the plane and the LDS plane is padded to K*NTOT so that no store
// needs a guard.
template <class TL, int NP, int NS, int NTOT>
struct ZStage {
  static constexpr int K0 = TL::PLANE / NTOT;
  static constexpr int R = TL::PLANE - K0 * NTOT;
  static constexpr bool PACK = R > 0 && K0 > 0 &&
                               (NP == 0 ? NS * R <= NTOT : NP * R <= 64 && NS * R <= 64 && NTOT >= 128);
  static constexpr int K = PACK ? K0 : (TL::PLANE + NTOT - 1) / NTOT;
  static constexpr int VS = PACK ? TL::PLANE : K * NTOT;
};

// LDS layout of a staged plane: NP pair planes, two variables interleaved
// cell by cell, then NS single planes (PAIRS), or one single plane per
// variable (!PAIRS).  One 128-bit LDS read (FP64) fetches
// one stencil tap of both variables of a pair: on gfx950 such reads move
// about 210 bytes per clock and CU, 64-bit reads of consecutive doubles
// about 120, and 64-bit reads of one half of each pair (a 16-byte stride)
// about 60 with bank conflicts.  So the variables without a partner get
// planes of their own, and a tap that only one component of a pair needs is
// still read at full width.  Pairs: (uu_x, uu_y), (uu_z, lnrho), [(aa_x,
// aa_y)]; singles: [aa_z], [ss].  The pair layout needs a few more registers
// (four per tap in flight instead of two); variants that would spill with it
// use single planes (zs_pairs).
template <class T>
struct alignas(2 * sizeof(T)) Pair {
  T x, y;
};

template <int C, class T>
__device__ __forceinline__ T pick(const Pair<T>& e) {
  if constexpr (C == 0) return e.x;
  else return e.y;
}
template <int C, class T>
__device__ __forceinline__ T pick(const T& e) {
  return e;
}

template <bool ENT, bool MAG, bool PAIRS>
struct ZLayout {
  using S = Slots<ENT, MAG>;
  static constexpr int NP = PAIRS ? (MAG ? 3 : 2) : 0;                 // pair planes
  static constexpr int NS = PAIRS ? (MAG ? 1 : 0) + (ENT ? 1 : 0) : S::NV;  // single planes
  static constexpr int NPL = NP + NS;                                  // staged planes
  static constexpr int WIDTH = 2 * NP + NS;                            // elements per cell
  static_assert(WIDTH == S::NV, "every variable staged once");
  // single planes of aa_z (MAG) and ss (ENT)
  static constexpr int SAZ = PAIRS ? 0 : S::IAX + 2;
  static constexpr int SSS = PAIRS ? (MAG ? 1 : 0) : S::ISS;
  // the variables of staged plane pl: components 0 and 1 of pair pl, or
  // (twice) the variable of single plane pl - NP
  __host__ __device__ static constexpr int v0(int pl) {
    if (!PAIRS) return pl;
    return pl == 0 ? 0 : pl == 1 ? 2 : pl < NP ? S::IAX : (MAG && pl == NP + SAZ ? S::IAX + 2 : S::ISS);
  }
  __host__ __device__ static constexpr int v1(int pl) {
    if (!PAIRS) return pl;
    return pl == 0 ? 1 : pl == 1 ? 3 : pl < NP ? S::IAX + 1 : v0(pl);
  }
};

// Layout choice per kernel variant (ENS: the ensemble kernel): pair planes
// unless the variant is known to run out of registers with them.  The
// deferred first-substep stage holds the state and the 2N register of the
// next plane in flight, and the ensemble kernel keeps each box's constants
// in registers; both spill with pairs in FP64.
template <typename T, bool MAG, bool ENT, bool POLY, int ST, class TL, bool ENS>
constexpr bool zs_pairs() {
  if constexpr (sizeof(T) == 4) return true;
  else if constexpr (ENS || ST == ST_S2D) return false;
  else if constexpr (std::is_same_v<TL, ZTile<16, 16>>)
    return !(ST == ST_S3 && !ENT && !POLY) && !(POLY && ST == ST_S2);
  else return !POLY && !(ENT && ST == ST_W);
}

template <int I>
using IC = std::integral_constant<int, I>;

// Ties v to sink without an instruction: the value counts as used, so the
// load that produced it keeps its full width.
template <class T>
__device__ __forceinline__ void hold_reg(T& sink, T v) {
  asm("" : "+v"(sink) : "v"(v));
}

// Values per cell the field half passes to the flow half: J x B / rho, then
// J^2 (entropy) and the Alfven Courant term (stages that reduce dt).
template <bool MAG, bool ENT, int ST>
constexpr int zs_nxc() {
  return MAG ? 3 + (ENT ? 1 : 0) + (StageT<ST>::DT ? 1 : 0) : 0;
}

// Ring slots.  Four hold planes q..q+3; the refill with plane q+4 must then
// wait until every wave has read plane q, a second barrier per plane.  With
// a fifth slot, plane q+4 goes into the slot of plane q-1 before the one
// barrier, and a double-buffered exchange lets the flow half finish plane q
// while the field half starts plane q+1.  Five slots where they fit in LDS
// (with the magnetic field: one block per CU in any case).
constexpr int ZS_LDS_BYTES = 163840;
template <typename T, bool MAG, bool ENT, int ST, class TL, bool PAIRS>
constexpr int zs_ring_slots() {
  if constexpr (!MAG) {
    return 4;
  } else {
    using LY = ZLayout<ENT, MAG, PAIRS>;
    constexpr long slot = (long)LY::WIDTH * ZStage<TL, LY::NP, LY::NS, zs_threads<MAG>()>::VS * sizeof(T);
    constexpr long xch2 = 2L * zs_nxc<MAG, ENT, ST>() * TL::NT * sizeof(T);
    constexpr long red = (long)(zs_threads<MAG>() / 64) * NRED * sizeof(double);
    return 5 * slot + xch2 + red <= ZS_LDS_BYTES ? 5 : 4;
  }
}

// Elements (of the state type) of the LDS ring and of the exchange buffer.
template <typename T, bool MAG, bool ENT, class TL, bool PAIRS, int RING>
constexpr int zs_ring_elems() {
  using LY = ZLayout<ENT, MAG, PAIRS>;
  return RING * LY::WIDTH * ZStage<TL, LY::NP, LY::NS, zs_threads<MAG>()>::VS;
}
template <bool MAG, bool ENT, int ST, class TL, int RING>
constexpr int zs_xch_elems() {
  return MAG ? (RING == 5 ? 2 : 1) * zs_nxc<MAG, ENT, ST>() * TL::NT : 1;
}

// Marches one tile column (x0, y0, z0 .. z0+lz-1) of one box through the
// RK stage ST.  ring and xch are the block's LDS buffers; the column's
// reduction terms are added to red (all threads of the block take part).
template <typename T, bool MAG, bool ENT, bool POLY, int ST, class TL, bool PAIRS, int RING, int UPW = 0,
          bool LAY = false>
__device__ __forceinline__ void zsplit_column(const KArgs& a, const Phys& P, int x0, int y0, int z0, int lz,
                                              T* ring, T* xch, double (&red)[NRED]) {
  using S = Slots<ENT, MAG>;
  using N = ZNeed<MAG, ENT>;
  using G = StageT<ST>;
  constexpr bool DT = G::DT;
  constexpr int NV = S::NV, IA = S::IAX, ISS = S::ISS;
  constexpr int NTOT = zs_threads<MAG>();
  using LY = ZLayout<ENT, MAG, PAIRS>;
  using T2 = Pair<T>;
  constexpr int NP = LY::NP, NPL = LY::NPL;
  using ZS = ZStage<TL, NP, LY::NS, NTOT>;
  constexpr int VS = ZS::VS;               // cells per staged plane (padded)
  constexpr int SLOT = LY::WIDTH * VS;     // elements per ring slot
  static_assert(VS % 2 == 0, "pair planes stay 16-byte aligned");
  // element offset in a ring slot of pair plane j / single plane s
  auto pair_at = [](const T* slot, int j) { return reinterpret_cast<const T2*>(slot + 2 * j * VS); };
  auto single_at = [](const T* slot, int sp) { return slot + (2 * NP + sp) * VS; };
  constexpr int NF = ENT ? 5 : 4;              // flow variables: uu, lnrho, [ss]
  constexpr int NGV = NF;                      // largest group (the flow half)
  constexpr bool MIXED = sizeof(T) == 4;
  constexpr int PX = TL::PX;

  const int tid = threadIdx.x;
  const int grp = MAG ? tid / TL::NT : 0;  // wave-uniform
  const int t = tid % TL::NT;
  const int tx = t % TL::TX, ty = t / TL::TX;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const long nxy = (long)nx * ny, nxyz = a.nxyz;
  const int zlast = z0 + lz - 1;

  auto zoff = [&](int p) -> long {
    int zz = p % nz;
    zz += (zz < 0) ? nz : 0;
    return (long)zz * nxy;
  };
  static_assert(RING == 4 || RING == 5, "four or five ring slots");
  auto slot_of = [&](int p) -> int { return RING == 4 ? (p - z0 + 3) & 3 : (p - z0 + 3) % 5; };
  // exchange buffer: values jxbr[0..2], [j2], [advec_va2]; two buffers
  // (by plane parity) with five ring slots
  constexpr int NXC = zs_nxc<MAG, ENT, ST>();
  constexpr int XJ2 = 3, XVA = 3 + (ENT ? 1 : 0);
  auto xbuf = [&](int q) -> T* { return xch + (RING == 5 ? (q & 1) * NXC * TL::NT : 0); };
  // offset inside a z plane of element ee of the tile's haloed plane
  auto plane_off = [&](int ee) -> unsigned {
    const int row = ee / PX, col = ee % PX;
    int gx = x0 - 3 + col, gy = y0 - 3 + row;
    gx += (gx < 0) ? nx : 0;
    gx -= (gx >= nx) ? nx : 0;
    gy += (gy < 0) ? ny : 0;
    gy -= (gy >= ny) ? ny : 0;
    return (unsigned)(gy * nx + gx);
  };

  // Plane staging (see ZStage): K per-variable slots plus, with PACK, the
  // shared remainder slot.
  constexpr int K = ZS::K;
  constexpr int R = ZS::R;
  unsigned poff[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int e = tid + k * NTOT;
    poff[k] = plane_off(e < TL::PLANE ? e : TL::PLANE - 1);  // padding lanes load a valid duplicate
  }
  // with pair planes, wave 0 stages pair remainders and wave 1 single ones
  const bool rem_pair = NP > 0 && tid < 64;
  const int rt = NP == 0 || rem_pair ? tid : tid - 64;
  const bool has_rem = ZS::PACK && (rem_pair ? rt < NP * R : rt < LY::NS * R);
  const int rp = has_rem ? (rem_pair ? 0 : NP) + rt / R : 0;  // staged plane of the remainder element
  const int rva = LY::v0(rp), rvb = LY::v1(rp);
  const int re = K * NTOT + (has_rem ? rt % R : 0);
  const unsigned roff = plane_off(re < TL::PLANE ? re : TL::PLANE - 1);
  // remainder element offset in a slot: in pairs for a pair plane, in
  // elements for a single plane
  const int rl = rp < NP ? rp * VS + re : (2 * NP + rp - NP) * VS + re;
  const double rmua = MIXED ? a.means[rva] : 0.0, rmub = MIXED ? a.means[rvb] : 0.0;

  // dt*beta1 of the deferred first-substep update (stage S2)
  const double dtb1 = a.d_dt ? a.d_dt[0] * a.beta1 : a.dtb1;
  constexpr int NW = G::FORM1 ? NV * K : 1;
  double pre[NV * K], prem[2] = {0.0, 0.0}, prw[NW], prwm[2] = {0.0, 0.0};
  // staged planes beyond a non-periodic z boundary come from the ghost-plane
  // buffer (variable stride 6 planes); the deferred-update stage S2D is not
  // run with one
  auto load_plane = [&](int p) {
    const double* src = zsrc(a, 0, p);
    const long vs = zsrc(a, 1, p) - src;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      const double* sv = src + v * vs;
#pragma unroll
      for (int k = 0; k < K; ++k) pre[v * K + k] = sv[poff[k]];
    }
    if (has_rem) {
      prem[0] = src[rva * vs + roff];
      prem[1] = src[rvb * vs + roff];
    }
    if constexpr (G::FORM1) {
      const double* srw = a.w + zoff(p);
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        const double* sv = srw + v * nxyz;
#pragma unroll
        for (int k = 0; k < K; ++k) prw[v * K + k] = sv[poff[k]];
      }
      if (has_rem) {
        prwm[0] = srw[rva * nxyz + roff];
        prwm[1] = srw[rvb * nxyz + roff];
      }
    }
  };
  // the staged value: f (or f1 = f0 + dt*beta1*w1 in stage S2)
  auto staged = [&](double f, double w) -> double {
    if constexpr (G::FORM1) return f + dtb1 * w;
    else return f;
  };
  auto to_t = [&](double val, double mu) -> T { return MIXED ? (T)(val - mu) : (T)val; };
  auto store_plane = [&](int p) {
    T* dst = ring + slot_of(p) * SLOT;
    T2* dst2 = reinterpret_cast<T2*>(dst);
#pragma unroll
    for (int pl = 0; pl < NPL; ++pl) {
      const int va = LY::v0(pl), vb = LY::v1(pl);
      const double mua = MIXED ? a.means[va] : 0.0;
      const double mub = MIXED ? a.means[vb] : 0.0;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const T ea = to_t(staged(pre[va * K + k], prw[G::FORM1 ? va * K + k : 0]), mua);
        if (pl < NP) {
          T2 e;
          e.x = ea;
          e.y = to_t(staged(pre[vb * K + k], prw[G::FORM1 ? vb * K + k : 0]), mub);
          dst2[pl * VS + tid + k * NTOT] = e;
        } else {
          dst[(2 * NP + pl - NP) * VS + tid + k * NTOT] = ea;
        }
      }
    }
    if (has_rem) {
      const T ea = to_t(staged(prem[0], prwm[0]), rmua);
      if (rem_pair) {
        T2 e;
        e.x = ea;
        e.y = to_t(staged(prem[1], prwm[1]), rmub);
        dst2[rl] = e;
      } else {
        dst[rl] = ea;
      }
    }
  };

  // the variables of this thread's group: flow vi -> 0,1,2,3,[ISS]; field vi -> IA+vi
  auto var_of = [&](int g, int i) -> int { return g == 0 ? (i < 4 ? i : ISS) : IA + i; };
  const int x = x0 + tx, y = y0 + ty;
  const unsigned own = (unsigned)(y * nx + x);  // offset inside a z plane
  const int c = (ty + 3) * PX + (tx + 3);
  // start of variable v's z plane p in an array of the state layout (uniform)
  auto vplane = [&](auto* base, int v, int p) { return base + ((long)v * nxyz + (long)p * nxy); };

  // own-cell prefetch of the next output plane: 2N register (READW) and, in
  // the mixed form, the FP64 state values (the ring holds FP32 offsets)
  constexpr int NOWNF = MIXED ? (MAG ? 4 : 0) : 0;  // field half also needs u, lnrho in FP64
  constexpr int NOWNW = G::FORM1 ? NOWNF : 0;       // ... and, in stage S2, their w1
  double wnx[NGV + NOWNW], fnx[NGV + NOWNF];
  auto prefetch_own = [&](auto g_c, int p) {
    constexpr int g = decltype(g_c)::value;
    constexpr int ngv = g == 0 ? NF : 3;
    p = p < z0 ? z0 : (p > zlast ? zlast : p);  // stays inside this block's chunk
    if constexpr (G::READW) {
#pragma unroll
      for (int i = 0; i < ngv; ++i) wnx[i] = __builtin_nontemporal_load(&vplane(a.w, var_of(g, i), p)[own]);
    }
    if constexpr (MIXED) {
#pragma unroll
      for (int i = 0; i < ngv; ++i) fnx[i] = vplane(a.fin, var_of(g, i), p)[own];
      if constexpr (MAG && g == 1) {
#pragma unroll
        for (int i = 0; i < 4; ++i) fnx[NGV + i] = vplane(a.fin, i, p)[own];
        if constexpr (G::FORM1) {
#pragma unroll
          for (int i = 0; i < 4; ++i) wnx[NGV + i] = __builtin_nontemporal_load(&vplane(a.w, i, p)[own]);
        }
      }
    }
  };

  // ---- prime the ring with planes z0-3 .. z0+RING-4, start the next ----
  for (int p = z0 - 3; p <= z0 + RING - 4; ++p) {
    load_plane(p);
    store_plane(p);
  }

  // carried state: z window of the own column, minus parts of the z-mixed
  // derivatives (group-local slots: xz of uu_x|aa_x -> 0, uu_z|aa_z -> 1;
  // yz of uu_y|aa_y -> 0, uu_z|aa_z -> 1)
  T back[NGV][3], cxz[2][3], cyz[2][3];
#pragma unroll
  for (int i = 0; i < NGV; ++i) back[i][0] = back[i][1] = back[i][2] = (T)0;
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    cxz[i][0] = cxz[i][1] = cxz[i][2] = (T)0;
    cyz[i][0] = cyz[i][1] = cyz[i][2] = (T)0;
  }
  const T fx1 = (T)P.fac1[0], fy1 = (T)P.fac1[1], fz1 = (T)P.fac1[2];
  const T fx2 = (T)P.fac2[0], fy2 = (T)P.fac2[1], fz2 = (T)P.fac2[2];
  const T C0 = (T)PCMX_D2C0, C1 = (T)PCMX_D2C1, C2 = (T)PCMX_D2C2, C3 = (T)PCMX_D2C3;
  const double dtb = a.d_dt ? a.d_dt[0] * a.beta : a.dtb;

  // |u| of the own cell for the upwind corrections of the flow half (set at
  // the start of each plane when UPW != 0)
  T ua[3] = {(T)0, (T)0, (T)0};
  // does variable V take an upwind correction?
  auto upwinded = [](auto v_c) {
    constexpr int V = decltype(v_c)::value;
    return std::bool_constant<(V < 3 && (UPW & UPW_UU) != 0) || (V == 3 && (UPW & UPW_LNRHO) != 0) ||
                              (ENT && V == ISS && (UPW & UPW_SS) != 0)>{};
  };

  // Stencils at plane q of the variables of one staged plane (PT = T2: a
  // pair, components 0 and 1 are variables VA and VB with group-local
  // indices IA_ and IB_; PT = T: a single plane, VB = -1).  Every tap is one
  // read of the plane; each component's arithmetic is that of a single
  // variable.  The taps are taken stage by stage (z, xz, yz, x, y, xy), both
  // components at once, so that each read is dead after its stage; a tap of
  // a pair that one component needs keeps the other half alive through
  // sink, which holds the read at full width.  Carries are updated with
  // plane q.  Returns the own-cell value(s).
  auto derive = [&](auto emit_c, auto va_c, auto ia_c, auto vb_c, auto ib_c, Derivs<T>& d, T& sink,
                    const auto* p0, const auto* p1, const auto* p2, const auto* p3) {
    using PT = std::remove_cv_t<std::remove_pointer_t<decltype(p0)>>;
    constexpr bool PAIR = std::is_same_v<PT, T2>;
    constexpr bool EMIT = decltype(emit_c)::value;
    constexpr int VA = decltype(va_c)::value, VB = decltype(vb_c)::value;
    static_assert(PAIR == (VB >= 0), "pairs evaluate both components");
    constexpr bool XY = N::xy(VA) || (VB >= 0 && N::xy(VB));
    constexpr bool XZ = N::xz(VA) || (VB >= 0 && N::xz(VB));
    constexpr bool YZ = N::yz(VA) || (VB >= 0 && N::yz(VB));
    // f(c_c, v_c, i_c) for each component
    auto each = [&](auto f) {
      f(IC<0>{}, va_c, ia_c);
      if constexpr (VB >= 0) f(IC<1>{}, vb_c, ib_c);
    };
    // keeps the unused half of pair taps that only component C uses (on the
    // (uu_z, lnrho) pair only: on the others the extra live registers of the
    // held halves would spill)
    auto hold = [&](auto c_c, auto need, auto... e) {
      constexpr int C = decltype(c_c)::value;
      if constexpr (PAIR && decltype(need)::value && VA == 2) {
        (hold_reg(sink, pick<1 - C>(e)), ...);
      }
    };
    const PT c0 = p0[0];
    T rdx[2] = {}, rdy[2] = {}, rdz[2] = {}, rxx[2] = {}, ryy[2] = {}, rzz[2] = {};
    T rxy[2] = {}, rxz[2] = {}, ryz[2] = {};
    T upx[2] = {}, upy[2] = {}, upz[2] = {};  // |u_j| D6_j f per direction
    if constexpr (EMIT) {
      // own column: z derivatives (der_main, der2_main; deriv.f90:124-162, 497-550)
      const PT zp1 = p1[0], zp2 = p2[0], zp3 = p3[0];
      each([&](auto c_c, auto v_c, auto i_c) {
        constexpr int C = decltype(c_c)::value, V = decltype(v_c)::value, I = decltype(i_c)::value;
        const T zm1 = back[I][0], zm2 = back[I][1], zm3 = back[I][2];
        rdz[C] = fz1 * ((T)45.0 * (pick<C>(zp1) - zm1) - (T)9.0 * (pick<C>(zp2) - zm2) + (pick<C>(zp3) - zm3));
        if constexpr (N::d2(V))
          rzz[C] = fz2 * (C0 * pick<C>(c0) + C1 * (pick<C>(zp1) + zm1) + C2 * (pick<C>(zp2) + zm2) +
                          C3 * (pick<C>(zp3) + zm3));
        if constexpr (decltype(upwinded(v_c))::value)
          upz[C] = ua[2] * der6_upwind<T>(fz1, pick<C>(c0), pick<C>(zp1) + zm1, pick<C>(zp2) + zm2,
                                          pick<C>(zp3) + zm3);
      });
      if constexpr (XZ) {  // forward half of the xz bidiagonal
        const PT a1 = p1[1], b1 = p1[-1], a2 = p2[2], b2 = p2[-2], a3 = p3[3], b3 = p3[-3];
        constexpr bool ONE = !(N::xz(VA) && VB >= 0 && N::xz(VB));
        hold(IC<N::xz(VA) ? 0 : 1>{}, std::bool_constant<ONE>{}, a1, b1, a2, b2, a3, b3);
        each([&](auto c_c, auto v_c, auto) {
          constexpr int C = decltype(c_c)::value, V = decltype(v_c)::value;
          if constexpr (N::xz(V)) {
            const T plus = (T)270.0 * (pick<C>(a1) - pick<C>(b1)) - (T)27.0 * (pick<C>(a2) - pick<C>(b2)) +
                           (T)2.0 * (pick<C>(a3) - pick<C>(b3));
            rxz[C] = (T)P.facxz * (plus - cxz[N::ixz(V) & 1][0]);
          }
        });
      }
      if constexpr (YZ) {  // forward half of the yz bidiagonal
        const PT a1 = p1[PX], b1 = p1[-PX], a2 = p2[2 * PX], b2 = p2[-2 * PX], a3 = p3[3 * PX], b3 = p3[-3 * PX];
        constexpr bool ONE = !(N::yz(VA) && VB >= 0 && N::yz(VB));
        hold(IC<N::yz(VA) ? 0 : 1>{}, std::bool_constant<ONE>{}, a1, b1, a2, b2, a3, b3);
        each([&](auto c_c, auto v_c, auto) {
          constexpr int C = decltype(c_c)::value, V = decltype(v_c)::value;
          if constexpr (N::yz(V)) {
            const T plus = (T)270.0 * (pick<C>(a1) - pick<C>(b1)) - (T)27.0 * (pick<C>(a2) - pick<C>(b2)) +
                           (T)2.0 * (pick<C>(a3) - pick<C>(b3));
            ryz[C] = (T)P.facyz * (plus - cyz[N::iyz(V) & 1][0]);
          }
        });
      }
    }
    {  // x: derivatives, and the minus half of the xz bidiagonal carried to planes q+1..q+3
      const PT xp1 = p0[1], xm1 = p0[-1], xp2 = p0[2], xm2 = p0[-2], xp3 = p0[3], xm3 = p0[-3];
      each([&](auto c_c, auto v_c, auto) {
        constexpr int C = decltype(c_c)::value, V = decltype(v_c)::value;
        const T g1 = pick<C>(xp1) - pick<C>(xm1), g2 = pick<C>(xp2) - pick<C>(xm2), g3 = pick<C>(xp3) - pick<C>(xm3);
        if constexpr (EMIT) {
          rdx[C] = fx1 * ((T)45.0 * g1 - (T)9.0 * g2 + g3);
          if constexpr (N::d2(V))
            rxx[C] = fx2 * (C0 * pick<C>(c0) + C1 * (pick<C>(xp1) + pick<C>(xm1)) +
                            C2 * (pick<C>(xp2) + pick<C>(xm2)) + C3 * (pick<C>(xp3) + pick<C>(xm3)));
          if constexpr (decltype(upwinded(v_c))::value)
            upx[C] = ua[0] * der6_upwind<T>(fx1, pick<C>(c0), pick<C>(xp1) + pick<C>(xm1),
                                            pick<C>(xp2) + pick<C>(xm2), pick<C>(xp3) + pick<C>(xm3));
        }
        if constexpr (N::xz(V)) {
          T* m = cxz[N::ixz(V) & 1];
          m[0] = m[1] + (T)270.0 * g1;
          m[1] = m[2] - (T)27.0 * g2;
          m[2] = (T)2.0 * g3;
        }
      });
    }
    {  // y: derivatives, and the minus half of the yz bidiagonal
      const PT yp1 = p0[PX], ym1 = p0[-PX], yp2 = p0[2 * PX], ym2 = p0[-2 * PX], yp3 = p0[3 * PX], ym3 = p0[-3 * PX];
      each([&](auto c_c, auto v_c, auto) {
        constexpr int C = decltype(c_c)::value, V = decltype(v_c)::value;
        const T g1 = pick<C>(yp1) - pick<C>(ym1), g2 = pick<C>(yp2) - pick<C>(ym2), g3 = pick<C>(yp3) - pick<C>(ym3);
        if constexpr (EMIT) {
          rdy[C] = fy1 * ((T)45.0 * g1 - (T)9.0 * g2 + g3);
          if constexpr (N::d2(V))
            ryy[C] = fy2 * (C0 * pick<C>(c0) + C1 * (pick<C>(yp1) + pick<C>(ym1)) +
                            C2 * (pick<C>(yp2) + pick<C>(ym2)) + C3 * (pick<C>(yp3) + pick<C>(ym3)));
          if constexpr (decltype(upwinded(v_c))::value)
            upy[C] = ua[1] * der6_upwind<T>(fy1, pick<C>(c0), pick<C>(yp1) + pick<C>(ym1),
                                            pick<C>(yp2) + pick<C>(ym2), pick<C>(yp3) + pick<C>(ym3));
        }
        if constexpr (N::yz(V)) {
          T* m = cyz[N::iyz(V) & 1];
          m[0] = m[1] + (T)270.0 * g1;
          m[1] = m[2] - (T)27.0 * g2;
          m[2] = (T)2.0 * g3;
        }
      });
    }
    if constexpr (EMIT) {
      if constexpr (XY) {  // bidiagonal xy, operand order of deriv.f90:1410-1418
        T bk[2][3] = {};
#pragma unroll
        for (int k = 1; k <= 3; ++k) {
          const PT pp = p0[k + k * PX], mp = p0[-k + k * PX], mm = p0[-k - k * PX], pm = p0[k - k * PX];
          each([&](auto c_c, auto v_c, auto) {
            constexpr int C = decltype(c_c)::value, V = decltype(v_c)::value;
            if constexpr (N::xy(V)) bk[C][k - 1] = pick<C>(pp) - pick<C>(mp) + pick<C>(mm) - pick<C>(pm);
          });
        }
        each([&](auto c_c, auto v_c, auto) {
          constexpr int C = decltype(c_c)::value, V = decltype(v_c)::value;
          if constexpr (N::xy(V))
            rxy[C] = (T)P.facxy * ((T)270.0 * bk[C][0] - (T)27.0 * bk[C][1] + (T)2.0 * bk[C][2]);
        });
      }
      each([&](auto c_c, auto v_c, auto) {
        constexpr int C = decltype(c_c)::value, V = decltype(v_c)::value;
        file_derivs<T, MAG, ENT>(d, V, rdx[C], rdy[C], rdz[C], rxx[C], ryy[C], rzz[C], rxy[C], rxz[C], ryz[C]);
        if constexpr (decltype(upwinded(v_c))::value) d.upw[V] = upx[C] + upy[C] + upz[C];  // x, y, z in order
      });
    }
    // the own column's z window moves up one plane
    each([&](auto c_c, auto, auto i_c) {
      constexpr int C = decltype(c_c)::value, I = decltype(i_c)::value;
      back[I][2] = back[I][1];
      back[I][1] = back[I][0];
      back[I][0] = pick<C>(c0);
    });
    return c0;
  };

  // values kept by the flow half between its two phases
  double wf[5], ff[5];
  FlowKeep<T> keep;
  double u2c = 0.0;

  // refill, after the flow half's second phase: the next plane replaces
  // plane q, which every wave has read before the barrier of this
  // iteration.  Four slots: plane q+4, read from the next iteration on, so a
  // second barrier follows.  Five slots: plane q+5, first read two
  // iterations on, past the next barrier, so none is needed.  Then the
  // plane after it starts loading (past the last plane the block needs, the
  // last one is fetched again, an L2 hit, so that every iteration issues the
  // same memory operations).
  constexpr int AHEAD = RING;  // plane stored by iteration q: q + AHEAD
  auto refill = [&](int q) {
    store_plane(q + AHEAD);
    load_plane(q + AHEAD + 1 <= zlast + 3 ? q + AHEAD + 1 : zlast + 3);
  };

  auto iteration = [&](auto emit_c, auto g_c, int q) {
    constexpr bool EMIT = decltype(emit_c)::value;
    constexpr int GRP = decltype(g_c)::value;
    const T* L0 = ring + slot_of(q) * SLOT;  // ring slots of planes q .. q+3
    const T* L1 = ring + slot_of(q + 1) * SLOT;
    const T* L2 = ring + slot_of(q + 2) * SLOT;
    const T* L3 = ring + slot_of(q + 3) * SLOT;
    // own-cell pointers into pair plane j / single plane sp of the four slots
    auto dpair = [&](auto j_c, auto va_c, auto ia_c, auto vb_c, auto ib_c, Derivs<T>& d, T& sink) {
      constexpr int j = decltype(j_c)::value;
      return derive(emit_c, va_c, ia_c, vb_c, ib_c, d, sink, pair_at(L0, j) + c, pair_at(L1, j) + c,
                    pair_at(L2, j) + c, pair_at(L3, j) + c);
    };
    auto dsingle = [&](auto sp_c, auto v_c, auto i_c, Derivs<T>& d, T& sink) {
      constexpr int sp = decltype(sp_c)::value;
      return derive(emit_c, v_c, i_c, IC<-1>{}, IC<-1>{}, d, sink, single_at(L0, sp) + c, single_at(L1, sp) + c,
                    single_at(L2, sp) + c, single_at(L3, sp) + c);
    };
    T sink = (T)0;
    // ---------------- phase 1 ----------------
    double wcur[NGV + NOWNW], fcur[NGV + NOWNF];
#pragma unroll
    for (int i = 0; i < NGV + NOWNW; ++i) wcur[i] = wnx[i];
#pragma unroll
    for (int i = 0; i < NGV + NOWNF; ++i) fcur[i] = fnx[i];
    if constexpr (MIXED && G::FORM1) {
      // own-cell FP64 state of the mixed form: f1 = f0 + dt*beta1*w1
      constexpr int ngv = GRP == 0 ? NF : 3;
#pragma unroll
      for (int i = 0; i < ngv; ++i) fcur[i] = fcur[i] + dtb1 * wcur[i];
      if constexpr (GRP == 1) {
#pragma unroll
        for (int i = 0; i < NOWNW; ++i) fcur[NGV + i] = fcur[NGV + i] + dtb1 * wcur[NGV + i];
      }
    }
    if constexpr (GRP == 0) {
      Derivs<T> d;
      if constexpr (EMIT && UPW != 0) {
        // |u| of the own cell (the FP64 values in the mixed form, whose ring
        // holds offsets from the box means)
        if constexpr (MIXED) {
#pragma unroll
          for (int j = 0; j < 3; ++j) ua[j] = (T)fabs(fcur[j]);
        } else if constexpr (PAIRS) {
          const T2 u01 = pair_at(L0, 0)[c], u2l = pair_at(L0, 1)[c];
          ua[0] = fabs(u01.x), ua[1] = fabs(u01.y), ua[2] = fabs(u2l.x);
        } else {
#pragma unroll
          for (int j = 0; j < 3; ++j) ua[j] = fabs(single_at(L0, j)[c]);
        }
      }
      T cv[5];
      if constexpr (PAIRS) {
        const T2 c01 = dpair(IC<0>{}, IC<0>{}, IC<0>{}, IC<1>{}, IC<1>{}, d, sink);
        const T2 c23 = dpair(IC<1>{}, IC<2>{}, IC<2>{}, IC<3>{}, IC<3>{}, d, sink);
        cv[0] = c01.x, cv[1] = c01.y, cv[2] = c23.x, cv[3] = c23.y;
      } else {
        cv[0] = dsingle(IC<0>{}, IC<0>{}, IC<0>{}, d, sink);
        cv[1] = dsingle(IC<1>{}, IC<1>{}, IC<1>{}, d, sink);
        cv[2] = dsingle(IC<2>{}, IC<2>{}, IC<2>{}, d, sink);
        cv[3] = dsingle(IC<3>{}, IC<3>{}, IC<3>{}, d, sink);
      }
      if constexpr (ENT) cv[4] = dsingle(IC<LY::SSS>{}, IC<ISS>{}, IC<4>{}, d, sink);
      asm volatile("" ::"v"(sink));
      if constexpr (EMIT) {
#pragma unroll
        for (int i = 0; i < NF; ++i) ff[i] = MIXED ? fcur[i] : (double)cv[i];
#pragma unroll
        for (int j = 0; j < 3; ++j) d.u[j] = (T)ff[j];
        d.lnrho = (T)ff[3];
        if constexpr (ENT) d.ss = (T)ff[4];
#pragma unroll
        for (int i = 0; i < NF; ++i) wf[i] = G::READW ? a.alpha * wcur[i] : 0.0;
        const T rho1 = pexp<T>(-d.lnrho);
        flow_point_pre<T, ENT, DT, POLY, UPW, LAY>(d, rho1, P, wf, keep, &u2c, LAY ? P.zlay + 4L * q : nullptr);
      }
    } else if constexpr (MAG) {
      Derivs<T> d;
      double fa[3];
      T cv[3];
      if constexpr (PAIRS) {
        const T2 cxy = dpair(IC<2>{}, IC<IA>{}, IC<0>{}, IC<IA + 1>{}, IC<1>{}, d, sink);
        cv[0] = cxy.x, cv[1] = cxy.y;
      } else {
        cv[0] = dsingle(IC<IA>{}, IC<IA>{}, IC<0>{}, d, sink);
        cv[1] = dsingle(IC<IA + 1>{}, IC<IA + 1>{}, IC<1>{}, d, sink);
      }
      cv[2] = dsingle(IC<LY::SAZ>{}, IC<IA + 2>{}, IC<2>{}, d, sink);
      asm volatile("" ::"v"(sink));
      if constexpr (EMIT) {
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          fa[i] = MIXED ? fcur[i] : (double)cv[i];
          d.a[i] = (T)fa[i];
        }
        T lnrho;
        if constexpr (MIXED) {
#pragma unroll
          for (int j = 0; j < 3; ++j) d.u[j] = (T)fcur[NGV + j];
          lnrho = (T)fcur[NGV + 3];
        } else {
          if constexpr (PAIRS) {
            const T2 u01 = pair_at(L0, 0)[c], u2l = pair_at(L0, 1)[c];
            d.u[0] = u01.x;
            d.u[1] = u01.y;
            d.u[2] = u2l.x;
            lnrho = u2l.y;
          } else {
#pragma unroll
            for (int j = 0; j < 3; ++j) d.u[j] = single_at(L0, j)[c];
            lnrho = single_at(L0, 3)[c];
          }
        }
        const T rho1 = pexp<T>(-lnrho);
        double wa[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) wa[i] = G::READW ? a.alpha * wcur[i] : 0.0;
        FieldOut<T> fo;
        double b2 = 0.0;
        field_point<T, ENT, DT>(d, rho1, P, wa, fo, &b2);
        T* xo = xbuf(q);
#pragma unroll
        for (int j = 0; j < 3; ++j) xo[j * TL::NT + t] = fo.jxbr[j];
        if constexpr (ENT) xo[XJ2 * TL::NT + t] = fo.j2;
        if constexpr (DT) xo[XVA * TL::NT + t] = fo.advec_va2;
        if constexpr (G::UPDATE) {
          double fo3[3];
#pragma unroll
          for (int i = 0; i < 3; ++i) fo3[i] = fa[i] + dtb * wa[i];
          if constexpr (ST == ST_S3 && !ENT)
            if (a.ftab && a.ftarget == IA)
              force_term(a.ftab, a.ffact, nx, a.ny, a.nz, x, y, q, fo3);
#pragma unroll
          for (int i = 0; i < 3; ++i) __builtin_nontemporal_store(fo3[i], &vplane(a.fout, IA + i, q)[own]);
        }
        if constexpr (G::WRITEW) {
#pragma unroll
          for (int i = 0; i < 3; ++i) __builtin_nontemporal_store(wa[i], &vplane(a.wout, IA + i, q)[own]);
        }
        if constexpr (G::DIAG) {
          red[2] += b2;
#pragma unroll
          for (int i = 0; i < 3; ++i) red[4 + IA + i] += fa[i];
        }
      }
    }
    prefetch_own(g_c, q + 1);
    __syncthreads();
    // ---------------- phase 2 ----------------
    if constexpr (EMIT) {
      if constexpr (GRP == 0) {
        FieldOut<T> fo{};
        if constexpr (MAG) {
          const T* xi = xbuf(q);
#pragma unroll
          for (int j = 0; j < 3; ++j) fo.jxbr[j] = xi[j * TL::NT + t];
          if constexpr (ENT) fo.j2 = xi[XJ2 * TL::NT + t];
          if constexpr (DT) fo.advec_va2 = xi[XVA * TL::NT + t];
        }
        double dt1 = 0.0;
        flow_point_post<T, MAG, ENT, DT, LAY>(keep, fo, P, wf, &dt1);
        if constexpr (G::UPDATE) {
          double fo3[NF];
#pragma unroll
          for (int i = 0; i < NF; ++i) fo3[i] = ff[i] + dtb * wf[i];
          if constexpr (ST == ST_S3 && !ENT) {  // with entropy: a separate pass (registers)
            if (a.ftab && a.ftarget == 0) {
              double fu[3] = {fo3[0], fo3[1], fo3[2]};
              force_term(a.ftab, a.ffact, nx, a.ny, a.nz, x, y, q, fu);
#pragma unroll
              for (int i = 0; i < 3; ++i) fo3[i] = fu[i];
            }
          }
#pragma unroll
          for (int i = 0; i < NF; ++i) __builtin_nontemporal_store(fo3[i], &vplane(a.fout, var_of(0, i), q)[own]);
        }
        if constexpr (G::WRITEW) {
#pragma unroll
          for (int i = 0; i < NF; ++i) __builtin_nontemporal_store(wf[i], &vplane(a.wout, var_of(0, i), q)[own]);
        }
        if constexpr (G::DIAG) {
          red[0] += u2c;
          red[1] = fmax(red[1], u2c);
          if constexpr (DT) red[3] = fmax(red[3], dt1);
#pragma unroll
          for (int i = 0; i < NF; ++i) red[4 + var_of(0, i)] += ff[i];
        }
      }
    }
    refill(q);
    if constexpr (RING == 4) __syncthreads();
  };

  // Each wave group runs its own copy of the marching loop: inside a loop
  // every wave then issues one fixed sequence of memory operations, which
  // lets the compiler count outstanding loads exactly (a loop shared by both
  // groups, with group branches around unequal numbers of loads and stores,
  // ends in conservative full waits).  Both copies pass the same barriers.
  auto run = [&](auto g_c) {
    // same issue order as the loop body: own-cell loads, then the plane
    prefetch_own(g_c, z0);
    load_plane(z0 + RING - 3);
    __syncthreads();
    for (int q = z0 - 3; q < z0; ++q) iteration(std::false_type{}, g_c, q);
    for (int q = z0; q <= zlast; ++q) iteration(std::true_type{}, g_c, q);
  };
  if constexpr (MAG) {
    if (grp == 1) run(std::integral_constant<int, 1>{});
    else run(std::integral_constant<int, 0>{});
  } else {
    run(std::integral_constant<int, 0>{});
  }

}

// One box: block b marches tile column wid(b) (XCD-aware: consecutive work
// items, neighbouring tiles that share halo rims, go to the same XCD, whose
// L2 then serves the rims).
// LAY: the samples/conv-slab layer physics (Phys::zlay row of each plane).
template <typename T, bool MAG, bool ENT, bool POLY, int ST, class TL, int UPW = 0, bool LAY = false>
__global__ void __launch_bounds__(zs_threads<MAG>()) k_rhs_zsplit(KArgs a, Phys P, ZPlan zp) {
  constexpr bool PAIRS = zs_pairs<T, MAG, ENT, POLY, ST, TL, false>();
  constexpr int RING = zs_ring_slots<T, MAG, ENT, ST, TL, PAIRS>();
  __shared__ __attribute__((aligned(16))) T ring[zs_ring_elems<T, MAG, ENT, TL, PAIRS, RING>()];
  __shared__ T xch[zs_xch_elems<MAG, ENT, ST, TL, RING>()];
  const long b = blockIdx.x;
  const long wid = zp.xcd_remap ? (b % 8) * (zp.nblocks / 8) + b / 8 : b;
  const int txt = (int)(wid % zp.ntx);
  const long rest = wid / zp.ntx;
  const int tyt = (int)(rest % zp.nty);
  const int chunk = (int)(rest / zp.nty);
  double red[NRED];
#pragma unroll
  for (int q = 0; q < NRED; ++q) red[q] = 0.0;
  zsplit_column<T, MAG, ENT, POLY, ST, TL, PAIRS, RING, UPW, LAY>(a, P, txt * TL::TX, tyt * TL::TY, chunk * zp.lz,
                                                                 zp.lz, ring, xch, red);
  if constexpr (StageT<ST>::DIAG) {
    if (a.partials) block_reduce_store<zs_threads<MAG>()>(red, a.partials, (long)blockIdx.x);
  }
}

// Ensemble of boxes: per-box launch arguments and one tile column per work
// item, all in device memory.
struct EnsBox {
  KArgs a;  // this box's arrays, sizes and RK coefficients for the launch
  Phys P;   // this box's constants
};
struct EnsWork {
  int box, x0, y0, z0, lz;
};

// Persistent launch over the work items of every box: the blocks of XCD x
// take the x-th contiguous eighth of the work list in turn (items of one box
// are contiguous, neighbouring tiles first), so tiles that share halo rims
// run on one XCD.  partials[w] receives work item w's reduction terms.  Every
// block passes the same number of barriers per item, and the loop bounds are
// uniform, so all waves of a block leave the loop together.
template <typename T, bool MAG, bool ENT, bool POLY, int ST, class TL>
__global__ void __launch_bounds__(zs_threads<MAG>()) k_rhs_zsplit_ens(const EnsBox* boxes, const EnsWork* work,
                                                                      int nwork, double* partials) {
  constexpr bool PAIRS = zs_pairs<T, MAG, ENT, POLY, ST, TL, true>();
  constexpr int RING = zs_ring_slots<T, MAG, ENT, ST, TL, PAIRS>();
  __shared__ __attribute__((aligned(16))) T ring[zs_ring_elems<T, MAG, ENT, TL, PAIRS, RING>()];
  __shared__ T xch[zs_xch_elems<MAG, ENT, ST, TL, RING>()];
  const int nb = gridDim.x;
  // with a block count divisible by 8 every XCD holds nb/8 blocks: XCD x then
  // takes the x-th eighth of the list; otherwise plain striding
  const bool by_xcd = (nb % 8) == 0;
  const int xcd = by_xcd ? blockIdx.x % 8 : 0, rank = by_xcd ? blockIdx.x / 8 : blockIdx.x;
  const int stride = by_xcd ? nb / 8 : nb;
  const int w0 = by_xcd ? (int)((long)nwork * xcd / 8) : 0;
  const int w1 = by_xcd ? (int)((long)nwork * (xcd + 1) / 8) : nwork;
  for (int w = w0 + rank; w < w1; w += stride) {
    const EnsWork it = work[w];
    const KArgs a = boxes[it.box].a;
    const Phys P = boxes[it.box].P;
    double red[NRED];
#pragma unroll
    for (int q = 0; q < NRED; ++q) red[q] = 0.0;
    zsplit_column<T, MAG, ENT, POLY, ST, TL, PAIRS, RING>(a, P, it.x0, it.y0, it.z0, it.lz, ring, xch, red);
    if constexpr (StageT<ST>::DIAG) {
      if (partials) block_reduce_store<zs_threads<MAG>()>(red, partials, (long)w);
    }
    __syncthreads();  // the next item reuses the ring
  }
}

template <typename T, bool MAG, bool ENT, int ST, class TL, int UPW = 0, bool LAY = false>
void launch_zsplit(const KArgs& a, const Phys& P, const ZPlan& zp, hipStream_t s) {
  if constexpr (LAY) {  // entropy runs only (split_lay_ok)
    hipLaunchKernelGGL((k_rhs_zsplit<T, MAG, ENT, false, ST, TL, UPW, true>), dim3((unsigned)zp.nblocks),
                       dim3(zs_threads<MAG>()), 0, s, a, P, zp);
  } else if (!ENT && P.polytropic) {
    hipLaunchKernelGGL((k_rhs_zsplit<T, MAG, false, true, ST, TL, UPW>), dim3((unsigned)zp.nblocks),
                       dim3(zs_threads<MAG>()), 0, s, a, P, zp);
  } else {
    hipLaunchKernelGGL((k_rhs_zsplit<T, MAG, ENT, false, ST, TL, UPW>), dim3((unsigned)zp.nblocks),
                       dim3(zs_threads<MAG>()), 0, s, a, P, zp);
  }
}

}  // namespace pcmx
