// Device kernels of the elastic isotropic 3-D propagator for gfx950.
//
// Scheme (reference: elastic_iso_float/src/kernelsElaGpu_3D.cu):
//   3-level leapfrog on 9 fields (vx vy vz sxx syy szz sxz sxy syz), 8th-order
//   staggered derivatives, cosine sponge. The reference runs, per fine step,
//   one stepper launch (kernelsElaGpu_3D.cu:110-460), 7 injection launches, a
//   damping launch over both time levels (:1182-1226) and 7 recording launches.
//   Here one fine step is two stencil launches (velocity update, stress update;
//   they read only level n so they are independent) plus one injection and one
//   recording launch, with the sponge folded into the stencil:
//       n = w * (w * o + M D(c) [+ secondary source])
//   where every stored level has been damped exactly once. This equals the
//   reference's "step, inject, damp p0 and p1, swap" sequence, whose p0 after
//   the swap is the previous level damped a second time.
//
// Mapping to the CDNA4 execution model:
//   - a block is 64 (z, one wave row, coalesced) x BX (x) threads and owns a
//     64 x BX column tile that it marches along y over a chunk of planes;
//   - fields needing x or z derivatives are staged per plane in LDS with a
//     4-cell halo; y derivatives come from an 8-deep register queue per
//     thread, so each plane of each field is read from HBM once per tile;
//   - blocks are dealt round-robin to the 8 XCDs by the hardware; an optional
//     remap gives each XCD a contiguous range of tiles instead (measured
//     slower on the 328 x 472 x 468 grid, so off by default).
// No MFMA: the stencil is not a dense contraction; the kernels are HBM-bound.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ela {

constexpr int FAT = 4;
constexpr int TZ = 64;       // z extent of a tile = one wavefront
constexpr int NXCD = 8;      // MI355X: 8 XCDs, blocks dealt round-robin

enum { VX = 0, VY, VZ, SXX, SYY, SZZ, SXZ, SXY, SYZ };
enum { BX_ = 0, BY_, BZ_, L2M, LAM, MXZ, MXY, MYZ };

// Stepper modes.
enum { MODE_FWD = 0,       // n = w(w o + M D c)        stepFwdGpu_3D
       MODE_ADJ = 1,       // n = w(w o - M D c)        stepAdjGpu_3D
       MODE_BORN_ADJ = 2,  // n = w(w o - D (M c))      stepAdjGpuBorn_3D
       MODE_BORN_ADJ_PM = 3  // the same with M c read premultiplied (q) and M n written (qn)
};

// Optional work fused into a stepper launch (bit flags).
enum { F_SS = 1,     // add the secondary sources ssl/ssr (Born forward)
       F_ACC = 2,    // accumulate the new level into accl/accr (coarse-time interpolation)
       F_STORE = 4,  // write the step's derivative combinations to dst (gradient recomputation)
       F_GRAD = 8,   // gradient imaging: grad[m] += new level x stored derivatives drv
       F_MED3 = 16,  // MODE_FWD / MODE_ADJ: form the 8 staggered, scaled medium values from (rho, lambda, mu)
       F_Q16 = 32    // FP32 with F_STORE / F_GRAD: the derivative combinations as block-scaled int16 (dq16, dqs)
};

// Derivative combinations stored as int16 (F_Q16): one scale per block of
// DQ_BLOCK values, the block being one 128-byte FP32 line of a z row, i.e.
// one half of a tile row's wave (rows and tiles start on a line), so that
// the scale of element i is dqs[i / DQ_BLOCK] for the steppers and the point
// kernels alike. q = rint(x / s), s = max |x| / 32767 over the block; the
// value restored is q s, within s / 2 (2^-16 of the block's largest value).
constexpr int DQ_BLOCK = 32;

template <typename T>
struct StepArgs {
  T* o[9];              // level n-1 in, level n+1 out (in place)
  const T* c[9];        // level n
  const T* m[8];        // scaled medium
  const T* ssl[9];      // secondary sources (left / right coarse slices)
  const T* ssr[9];
  T* accl[9];           // coarse-time accumulation targets
  T* accr[9];
  T* dst[9];            // F_STORE: sx sy sz exx eyy ezz gxz gxy gyz of this step
  const T* drv[9];      // F_GRAD: the same combinations stored by the matching forward step
  T* grad[8];           // F_GRAD: gradient accumulators, one per medium array
  const T* q[9];        // MODE_BORN_ADJ_PM: M c of the current level
  T* qn[9];             // MODE_BORN_ADJ_PM: M n of the new level (written)
  const T* ep[3];       // F_MED3: rho, lambda, mu
  int16_t* dq16[9];     // F_Q16: int16 derivative combinations (written by F_STORE, read by F_GRAD)
  float* dqs[9];        // F_Q16: their scales, one per DQ_BLOCK elements
  T tdt;                // F_MED3: 2 dtw
  double tdt64;         // F_MED3: 2 dtw in FP64 (FP32 steppers form buoyancies and normal moduli in FP64)
  T ssfl, ssfr;         // filter weights for the secondary source
  T accfl, accfr;       // filter weights for the accumulation
  T cx[4], cy[4], cz[4];
  const T* taper;       // [min_pad]
  int min_pad;
  int nz, nx, ny;       // grid sizes (reference layout)
  int nzi, zoff;        // internal row pitch and z origin (line-aligned interior)
  int64_t ys;           // nx * nzi
  int ychunk;
  int nbz, nbx, nby;    // tile counts along z, x, y-chunks
  int remap;            // tile order (see tile_coord)
  int xgroup;           // x columns per group of the y-major order
  const int* chunks;    // stacked windows: per y chunk {y0, y1, window y0, window ny, window nx}, or null
};

__device__ __forceinline__ int xcd_remap(int bid, int nb) {
  const int per = nb / NXCD, rem = nb % NXCD;
  const int xcd = bid % NXCD, k = bid / NXCD;
  return xcd < rem ? xcd * (per + 1) + k : rem * (per + 1) + (xcd - rem) * per + k;
}

// Groups of `gs` consecutive tiles dealt whole to one XCD: the k-th block that
// XCD x receives runs tile ((k / gs) * NXCD + x) * gs + k % gs, so the tiles
// running at one time still form one contiguous range of the enumeration (as
// in the hardware order) but each group's halo lines meet in one L2. Blocks
// past the last full round of NXCD groups keep the identity mapping.
__device__ __forceinline__ int xcd_group(int bid, int nb, int gs) {
  const int full = nb / (NXCD * gs) * (NXCD * gs);
  if (bid >= full) return bid;
  const int x = bid % NXCD, k = bid / NXCD;
  return ((k / gs) * NXCD + x) * gs + k % gs;
}

// Sponge factor from the distance to the nearest face of the cell's grid (or
// of its window when several windows are stacked along y).
template <typename T>
__device__ __forceinline__ T damp_w(const StepArgs<T>& a, int iz, int ix, int iy, int wy0, int wny, int wnx) {
  int d = min(min(iz - FAT, ix - FAT), min(a.nz - iz - 1 - FAT, wnx - ix - 1 - FAT));
  d = min(d, min(iy - wy0 - FAT, wy0 + wny - iy - 1 - FAT));
  return d < a.min_pad ? a.taper[d] : T(1);
}

template <typename T>
__device__ __forceinline__ T ld(const T* __restrict__ p, int64_t i, bool ok) {
  return ok ? p[i] : T(0);
}

// y-derivative from an 8-deep queue q[j] = f[iy - s + j], s = 3 for D+ and 4
// for D-: both reduce to sum_k c_k (q[4+k] - q[3-k]).
template <typename T>
__device__ __forceinline__ T dq(const T* q, const T* c) {
  return c[0] * (q[4] - q[3]) + c[1] * (q[5] - q[2]) + c[2] * (q[6] - q[1]) + c[3] * (q[7] - q[0]);
}

// LDS plane with a 4-cell halo on both sides in x (rows) and z (columns).
template <typename T, int BX>
using Plane = T[BX + 2 * FAT][TZ + 2 * FAT];

// D+ / D- along z (columns) and x (rows) of an LDS plane, at (r, col) = (ty+4, tz+4).
template <typename T, int BX>
__device__ __forceinline__ T dpz(const Plane<T, BX>& p, int r, int q, const T* c) {
  return c[0] * (p[r][q + 1] - p[r][q]) + c[1] * (p[r][q + 2] - p[r][q - 1]) +
         c[2] * (p[r][q + 3] - p[r][q - 2]) + c[3] * (p[r][q + 4] - p[r][q - 3]);
}
template <typename T, int BX>
__device__ __forceinline__ T dmz(const Plane<T, BX>& p, int r, int q, const T* c) {
  return c[0] * (p[r][q] - p[r][q - 1]) + c[1] * (p[r][q + 1] - p[r][q - 2]) +
         c[2] * (p[r][q + 2] - p[r][q - 3]) + c[3] * (p[r][q + 3] - p[r][q - 4]);
}
template <typename T, int BX>
__device__ __forceinline__ T dpx(const Plane<T, BX>& p, int r, int q, const T* c) {
  return c[0] * (p[r + 1][q] - p[r][q]) + c[1] * (p[r + 2][q] - p[r - 1][q]) +
         c[2] * (p[r + 3][q] - p[r - 2][q]) + c[3] * (p[r + 4][q] - p[r - 3][q]);
}
template <typename T, int BX>
__device__ __forceinline__ T dmx(const Plane<T, BX>& p, int r, int q, const T* c) {
  return c[0] * (p[r][q] - p[r - 1][q]) + c[1] * (p[r + 1][q] - p[r - 2][q]) +
         c[2] * (p[r + 2][q] - p[r - 3][q]) + c[3] * (p[r + 3][q] - p[r - 4][q]);
}

// Guard past the end of every library array (unpredicated stencil loads).
constexpr int GUARD_ROWS = 16;      // >= the largest tile_x
constexpr int GUARD_EXTRA = 2 * TZ;  // >= one tile height plus its z halo

// 32-bit element offsets from a uniform base pointer: the compiler can use
// the scalar-base + 32-bit vector-offset addressing mode, which saves the
// 64-bit per-lane address arithmetic. The host guarantees that every array
// is smaller than 4 GiB (ela_create checks n * sizeof(T) < 2^32).
template <typename T>
__device__ __forceinline__ const T* at(const T* base, uint32_t i) {
  return reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + (uint64_t)(i * (uint32_t)sizeof(T)));
}
template <typename T>
__device__ __forceinline__ T* at(T* base, uint32_t i) {
  return reinterpret_cast<T*>(reinterpret_cast<char*>(base) + (uint64_t)(i * (uint32_t)sizeof(T)));
}
// Streaming accesses (read once / written once per step). The non-temporal
// cache policy is selectable at build time; measured on MI355X it slows the
// stencil down when applied to the loads, so loads default to the normal
// policy.
#ifndef ELA_NT_LOADS
#define ELA_NT_LOADS 0
#endif
#ifndef ELA_NT_STORES
#define ELA_NT_STORES 1
#endif
template <typename T>
__device__ __forceinline__ T ld_nt(const T* base, uint32_t i) {
#if ELA_NT_LOADS
  return __builtin_nontemporal_load(at(base, i));
#else
  return *at(base, i);
#endif
}
template <typename T>
__device__ __forceinline__ void st_nt(T* base, uint32_t i, T v) {
#if ELA_NT_STORES
  __builtin_nontemporal_store(v, at(base, i));
#else
  *at(base, i) = v;
#endif
}

// FP32 LDS rows loaded straight from HBM (1) or staged through registers (0).
#ifndef ELA_LDS_DMA
#define ELA_LDS_DMA 1
#endif

// Direct global -> LDS load of one row of a tile (FP32): lane l of the wave
// writes row[l], so a 64-lane wave fills TZ consecutive LDS values without
// passing them through registers. Completion is covered by the wait that
// precedes the block's next barrier.
__device__ __forceinline__ void row_to_lds(const float* base, uint32_t i, float* row) {
  __builtin_amdgcn_global_load_lds(at(base, i), (__attribute__((address_space(3))) void*)row, 4, 0, 0);
}

// Largest magnitude over each 32-lane half of the wave (one DQ_BLOCK), in
// every lane: DPP lane moves inside each 16-lane row (swap neighbours, swap
// pairs, mirror within 8, mirror within 16) and one swizzle that exchanges
// the two rows of a half (lane ^ 16). All 64 lanes must be active. The
// magnitudes are compared as unsigned integers (the bit patterns of
// non-negative floats order as their values), which needs no NaN quieting.
template <int CTRL>
__device__ __forceinline__ uint32_t dpp_umax(uint32_t u) {
  return max(u, (uint32_t)__builtin_amdgcn_mov_dpp((int)u, CTRL, 0xF, 0xF, true));
}

__device__ __forceinline__ float half_wave_absmax(float x) {
  uint32_t u = __float_as_uint(fabsf(x));
  u = dpp_umax<0xB1>(u);   // quad_perm [1, 0, 3, 2]
  u = dpp_umax<0x4E>(u);   // quad_perm [2, 3, 0, 1]
  u = dpp_umax<0x141>(u);  // row_half_mirror
  u = dpp_umax<0x140>(u);  // row_mirror
  u = max(u, (uint32_t)__builtin_amdgcn_ds_swizzle((int)u, 0x401F));  // lane ^ 16
  return __uint_as_float(u);
}

// F_Q16 store of one derivative combination at element i: every lane of the
// wave calls it; lanes outside the interior pass x = 0. The last z tile of a
// row may reach past the row into the next row's leading non-interior line
// ([0, zoff + FAT) = [0, 32)), which no stepper reads and no other tile
// writes. The hardware reciprocal (1 ulp) serves for 1 / m: |x| 32767 rcp(m)
// stays below 32767.5, so q fits int16, and the stored scale is what the
// reader multiplies by, so the error stays within s / 2 plus a few FP32 ulps.
__device__ __forceinline__ void store_dq16(int16_t* q, float* sc, uint32_t i, float x) {
  const float m = half_wave_absmax(x);
  const float inv = m > 0.f ? 32767.f * __builtin_amdgcn_rcpf(m) : 0.f;
  st_nt(q, i, (int16_t)__float2int_rn(x * inv));
  if ((threadIdx.x & (DQ_BLOCK - 1)) == 0) *at(sc, i / DQ_BLOCK) = m * (1.f / 32767.f);
}

__device__ __forceinline__ float load_dq16(const int16_t* q, const float* sc, uint32_t i) {
  return (float)*at(q, i) * *at(sc, i / DQ_BLOCK);
}

// Values that the derivative operators see. In MODE_BORN_ADJ they are the
// medium-multiplied fields (stepAdjGpuBorn_3D multiplies at every loaded
// position, kernelsElaGpu_3D.cu:864-1040); MODE_BORN_ADJ_PM reads the same
// products from the premultiplied level that the previous step wrote, so its
// halo and queue loads are one array per field as in the forward stepper.
//
// The loads are not predicated: a tile's columns, halo rows and halo column
// may lie outside the grid (past nz in the row, past nx in the plane), where
// no interior cell's stencil reaches, so those values are never used; every
// library array ends with a guard of GUARD_ROWS rows plus GUARD_EXTRA elements
// past its last plane, which holds the farthest such position (at most
// (tile_x - 1) rows plus one tile height past the array's end), so the loads
// stay inside the allocation.
template <typename T, int MODE>
struct StressVal {
  const StepArgs<T>& a;
  __device__ __forceinline__ T pm(int f, uint32_t i) const { return *at(a.q[f], i); }
  __device__ __forceinline__ T sxx(uint32_t i) const {
    if (MODE == MODE_BORN_ADJ_PM) return pm(SXX, i);
    if (MODE == MODE_BORN_ADJ)
      return *at(a.c[SXX], i) * *at(a.m[L2M], i) + (*at(a.c[SYY], i) + *at(a.c[SZZ], i)) * *at(a.m[LAM], i);
    return *at(a.c[SXX], i);
  }
  __device__ __forceinline__ T syy(uint32_t i) const {
    if (MODE == MODE_BORN_ADJ_PM) return pm(SYY, i);
    if (MODE == MODE_BORN_ADJ)
      return *at(a.c[SYY], i) * *at(a.m[L2M], i) + (*at(a.c[SXX], i) + *at(a.c[SZZ], i)) * *at(a.m[LAM], i);
    return *at(a.c[SYY], i);
  }
  __device__ __forceinline__ T szz(uint32_t i) const {
    if (MODE == MODE_BORN_ADJ_PM) return pm(SZZ, i);
    if (MODE == MODE_BORN_ADJ)
      return *at(a.c[SZZ], i) * *at(a.m[L2M], i) + (*at(a.c[SXX], i) + *at(a.c[SYY], i)) * *at(a.m[LAM], i);
    return *at(a.c[SZZ], i);
  }
  __device__ __forceinline__ T sxz(uint32_t i) const {
    if (MODE == MODE_BORN_ADJ_PM) return pm(SXZ, i);
    return MODE == MODE_BORN_ADJ ? *at(a.c[SXZ], i) * *at(a.m[MXZ], i) : *at(a.c[SXZ], i);
  }
  __device__ __forceinline__ T sxy(uint32_t i) const {
    if (MODE == MODE_BORN_ADJ_PM) return pm(SXY, i);
    return MODE == MODE_BORN_ADJ ? *at(a.c[SXY], i) * *at(a.m[MXY], i) : *at(a.c[SXY], i);
  }
  __device__ __forceinline__ T syz(uint32_t i) const {
    if (MODE == MODE_BORN_ADJ_PM) return pm(SYZ, i);
    return MODE == MODE_BORN_ADJ ? *at(a.c[SYZ], i) * *at(a.m[MYZ], i) : *at(a.c[SYZ], i);
  }
};

template <typename T, int MODE>
struct VelVal {
  const StepArgs<T>& a;
  __device__ __forceinline__ T v(int f, uint32_t i) const {
    if (MODE == MODE_BORN_ADJ_PM) return *at(a.q[f], i);
    return MODE == MODE_BORN_ADJ ? *at(a.c[f], i) * *at(a.m[f], i) : *at(a.c[f], i);  // BX_/BY_/BZ_ == VX/VY/VZ
  }
};

// Per-thread tile geometry, halo roles and 32-bit column offsets.
//
// Internal layout: rows of nzi elements (a multiple of one 128-byte cache
// line) with the reference's z = 0 stored at zoff, chosen so that the first
// interior cell (z = FAT) starts a cache line; a 64-wide tile then covers
// exactly two (FP32) lines and its z halos are the adjacent lines.
struct TileCoord {
  int tz, ty, iz, izo, ix, y0, y1, z0, x0;   // iz internal, izo reference z
  int wy0, wny, wnx;       // the tile's window: first plane, planes, x extent (whole grid without windows)
  bool empty;              // no interior cell in the tile (its window is narrower than the tile's x)
  bool okc, active;        // own column inside the grid / inside the interior
  bool xin;                // the tile row (wave-uniform) is an interior x row
  bool xh, okxl, okxh;     // x-halo loader (ty < 4) and its low/high validity
  bool zh, okzh;           // z-halo loader (lanes 0-7) and validity
  int colh;                // LDS column of the z-halo value
  uint32_t col, cxl, cxh, czh;
};

// Block -> tile. order 0: the hardware's dealing (block b runs on XCD b % 8)
// over tiles enumerated z fastest, then x, then y. order 1: each XCD gets a
// contiguous range of that enumeration. order 2: contiguous ranges of a
// y-major enumeration (per z tile, groups of `xgroup` x columns, y chunks
// within a group, columns fastest), so that the blocks an XCD runs at the
// same time are y- and x-neighbours whose halo planes meet in its L2.
// order 3: the z-fastest enumeration dealt to the XCDs in groups of whole z
// columns (`xgroup` adjacent columns of nbz tiles), so that z (and, with
// xgroup > 1, x) neighbours share an L2 while the grid is swept in the
// hardware order.
template <int BX>
__device__ __forceinline__ TileCoord tile_coord(int nbz, int nbx, int nby, int ychunk, int nz, int nzi, int zoff,
                                                int nx, int ny, int order, int xgroup, const int* chunks) {
  const int nb = nbz * nbx * nby;
  const int lb = order == 3 ? xcd_group(blockIdx.x, nb, nbz * xgroup)
                 : order ? xcd_remap(blockIdx.x, nb) : (int)blockIdx.x;
  int bz, bx, by;
  if (order == 2) {
    const int per_z = nbx * nby;
    bz = lb / per_z;
    const int p = lb - bz * per_z;
    const int grp = p / (xgroup * nby);
    const int r = p - grp * xgroup * nby;
    const int width = min(xgroup, nbx - grp * xgroup);
    by = r / width;
    bx = grp * xgroup + (r - by * width);
  } else {
    bz = lb % nbz;
    bx = (lb / nbz) % nbx;
    by = lb / (nbz * nbx);
  }
  TileCoord t;
  t.tz = threadIdx.x;
  t.ty = threadIdx.y;
  t.z0 = zoff + FAT + bz * TZ;
  t.x0 = FAT + bx * BX;
  t.iz = t.z0 + t.tz;
  t.izo = t.iz - zoff;
  t.ix = t.x0 + t.ty;
  if (chunks) {
    const int* c = chunks + 8 * by;
    t.y0 = c[0]; t.y1 = c[1]; t.wy0 = c[2]; t.wny = c[3]; t.wnx = c[4];
  } else {
    t.y0 = FAT + by * ychunk;
    t.y1 = min(t.y0 + ychunk, ny - FAT);
    t.wy0 = 0; t.wny = ny; t.wnx = nx;
  }
  nx = t.wnx;
  t.empty = t.x0 >= nx - FAT;
  t.okc = t.izo < nz && t.ix < nx;
  t.xin = t.ix < nx - FAT;
  t.active = (t.izo < nz - FAT) && t.xin;
  t.xh = t.ty < FAT;
  const int ixlo = t.x0 - FAT + t.ty, ixhi = t.x0 + BX + t.ty;
  t.okxl = t.xh && t.izo < nz;
  t.okxh = t.okxl && ixhi < nx;
  t.zh = t.tz < 2 * FAT;
  const int izh = t.tz < FAT ? t.z0 - FAT + t.tz : t.z0 + TZ + (t.tz - FAT);
  t.colh = t.tz < FAT ? t.tz : TZ + t.tz;
  t.okzh = t.zh && (izh - zoff) < nz && t.ix < nx;
  t.col = (uint32_t)(t.ix * nzi + t.iz);
  t.cxl = (uint32_t)(ixlo * nzi + t.iz);
  t.cxh = (uint32_t)(ixhi * nzi + t.iz);
  t.czh = (uint32_t)(t.ix * nzi + izh);
  return t;
}

// ---------------------------------------------------------------------------
// Velocity update: vx, vy, vz from the six stresses.
//
// Software pipeline per y plane: the global loads of plane iy+1 (new queue
// values, LDS centres and halos) are issued before plane iy is computed and
// land in the second LDS buffer, so one barrier per plane suffices and the
// HBM latency overlaps the stencil arithmetic.
// ---------------------------------------------------------------------------
// The velocity stepper's LDS: double-buffered planes of sxx, szz, sxy, sxz,
// syz and (F_MED3) density planes with the low x row and low z column.
template <typename T, int BX, bool MED3>
struct VelLds {
  Plane<T, BX> xx[2], zz[2], xy[2], xz[2], yz[2];
  T rho[2][MED3 ? BX + 1 : 1][MED3 ? TZ + 1 : 1];
};

template <typename T>
struct VelStage {
  T xx, zz, xz;                 // LDS centres of the non-queue fields
  T hxx0, hxx1, hxy0, hxy1, hxz0, hxz1;  // x-halo rows (low / high)
  T hzz, hxz, hyz;              // z-halo column
  T nxy, nyy, nyz;              // new queue values (plane + 4, + 3, + 4)
  T rho, rho_hx, rho_hz;        // F_MED3: density of the column, of row x0 - 1 and of z0 - 1
};

// F_MED3: the staggered, scaled medium formed in the stepper from (rho,
// lambda, mu) with the interior rule of the host scaling (a cell's staggered
// value averages it with its lower neighbour along the axis; the FD interior
// never meets the edge rules): rhox = (rho + rho[x-1]) / 2, rhoxDtw = 2 dtw /
// rhox (likewise y, z); lamb2MuDtw = 2 dtw (lambda + 2 mu), lambDtw = 2 dtw
// lambda; muxzDtw = 2 dtw * Sz(Sx(mu)) with S the same half-cell average, muxy
// = Sy(Sx(mu)), muyz = Sz(Sy(mu)) -- the operations of ela_scale_medium in the
// same order, so an FP64 stepper reproduces the 8-array stepper exactly.
//
// Buoyancy 2 dtw / rho_h of a staggered density rho_h. FP64: the host
// scaling's division. FP32: the same correctly rounded FP64 division of the
// FP64 half-cell sum of the FP32 densities, rounded once to FP32 -- the host's
// FP64 scaling of the same densities rounded to FP32, bit for bit. Formed in
// FP32 arithmetic, the buoyancies and moduli carried one or two extra
// roundings whose bias more than doubled the FP32 data's distance from FP64
// over a 5000-step run. k_scale_medium uses the same functions, so every mode
// of one handle sees the same medium.
template <typename T>
__device__ __forceinline__ T buoy_h(T tdt, double tdt64, double rho_h) {
  if constexpr (sizeof(T) == 4) {
    return T(tdt64 / rho_h);
  } else {
    return tdt / T(rho_h);
  }
}

template <typename T>
__device__ __forceinline__ T stag_buoy(const StepArgs<T>& a, T c, T m) {
  if constexpr (sizeof(T) == 4) return buoy_h<T>(a.tdt, a.tdt64, 0.5 * (double(c) + double(m)));
  else return a.tdt / (T(0.5) * (c + m));
}

// lamb2MuDtw and lambDtw of a cell (FP32: formed in FP64, rounded once).
template <typename T>
__device__ __forceinline__ void normal_moduli(T tdt, double tdt64, T lam, T mu, T& l2m, T& la) {
  if constexpr (sizeof(T) == 4) {
    l2m = T(tdt64 * (double(lam) + 2.0 * double(mu)));
    la = T(tdt64 * double(lam));
  } else {
    l2m = tdt * (lam + T(2) * mu);
    la = tdt * lam;
  }
}

template <typename T, int BX, int MODE, int F>
__global__ void __launch_bounds__(TZ * BX) k_velocity(StepArgs<T> a) {
  constexpr bool SS = (F & F_SS) != 0, ACC = (F & F_ACC) != 0, MED3 = (F & F_MED3) != 0;
  constexpr bool Q16 = (F & F_Q16) != 0, QS = Q16 && (F & F_STORE) != 0;
  static_assert(!Q16 || sizeof(T) == 4, "F_Q16 is an FP32 format");
  static_assert(!QS || (!SS && !ACC && (F & F_GRAD) == 0 && MODE == MODE_FWD), "F_Q16 stores: MODE_FWD recomputation");
  constexpr bool HOIST = sizeof(T) == 4;  // per-cell operands loaded ahead of the stores
  static_assert(!MED3 || MODE == MODE_FWD || MODE == MODE_ADJ || MODE == MODE_BORN_ADJ_PM,
                "F_MED3 needs MODE_FWD, MODE_ADJ or MODE_BORN_ADJ_PM");
  // one LDS object: the compiler's wait analysis then tells the rows the
  // direct loads fill (buffer b ^ 1) from the rows the stencil reads (buffer b)
  __shared__ VelLds<T, BX, MED3> L;
  auto& l_xx = L.xx; auto& l_zz = L.zz; auto& l_xy = L.xy; auto& l_xz = L.xz; auto& l_yz = L.yz;
  auto& l_rho = L.rho;
  const TileCoord t = tile_coord<BX>(a.nbz, a.nbx, a.nby, a.ychunk, a.nz, a.nzi, a.zoff, a.nx, a.ny, a.remap, a.xgroup, a.chunks);
  if (t.empty) return;  // block-uniform, before any barrier
  const uint32_t ys = (uint32_t)a.ys;
  const StressVal<T, MODE> S{a}, SXl{a}, SXh{a}, SZ{a};
  // FP32 rows that are plain array values go straight from HBM to LDS
  // (MODE_BORN_ADJ multiplies them by the medium first, in registers)
  constexpr bool DMA = ELA_LDS_DMA && sizeof(T) == 4 && MODE != MODE_BORN_ADJ;
  auto src = [&](int f) -> const T* { return MODE == MODE_BORN_ADJ_PM ? a.q[f] : a.c[f]; };
  const int r = t.ty + FAT, q = t.tz + FAT;
  // a wave is one row of the tile (blockDim.x == TZ == wavefront size): its row index is wave-uniform
  const int wy = __builtin_amdgcn_readfirstlane(t.ty), wr = wy + FAT;

  // plane p's values: LDS rows of buffer nb directly (DMA), the rest into v
  auto fetch = [&](VelStage<T>& v, int p, int nb) {
    const uint32_t pl = (uint32_t)p * ys;
    if constexpr (DMA) {
      row_to_lds(src(SXX), pl + t.col, &l_xx[nb][wr][FAT]);
      row_to_lds(src(SZZ), pl + t.col, &l_zz[nb][wr][FAT]);
      row_to_lds(src(SXZ), pl + t.col, &l_xz[nb][wr][FAT]);
    } else {
      v.xx = S.sxx(pl + t.col);
      v.zz = S.szz(pl + t.col);
      v.xz = S.sxz(pl + t.col);
    }
    v.nxy = S.sxy(pl + 4 * ys + t.col);
    v.nyy = S.syy(pl + 3 * ys + t.col);
    v.nyz = S.syz(pl + 4 * ys + t.col);
    if (wy < FAT) {  // the x-halo loaders (t.xh) are whole waves
      if constexpr (DMA) {
        row_to_lds(src(SXX), pl + t.cxl, &l_xx[nb][wy][FAT]); row_to_lds(src(SXX), pl + t.cxh, &l_xx[nb][wr + BX][FAT]);
        row_to_lds(src(SXY), pl + t.cxl, &l_xy[nb][wy][FAT]); row_to_lds(src(SXY), pl + t.cxh, &l_xy[nb][wr + BX][FAT]);
        row_to_lds(src(SXZ), pl + t.cxl, &l_xz[nb][wy][FAT]); row_to_lds(src(SXZ), pl + t.cxh, &l_xz[nb][wr + BX][FAT]);
      } else {
        v.hxx0 = SXl.sxx(pl + t.cxl); v.hxx1 = SXh.sxx(pl + t.cxh);
        v.hxy0 = SXl.sxy(pl + t.cxl); v.hxy1 = SXh.sxy(pl + t.cxh);
        v.hxz0 = SXl.sxz(pl + t.cxl); v.hxz1 = SXh.sxz(pl + t.cxh);
      }
    }
    if (t.zh) {
      v.hzz = SZ.szz(pl + t.czh); v.hxz = SZ.sxz(pl + t.czh); v.hyz = SZ.syz(pl + t.czh);
    }
    if (MED3) {
      // the column's own cell and its x - 1 / z - 1 neighbours
      if constexpr (DMA) {
        row_to_lds(a.ep[0], pl + t.col, &l_rho[nb][wy + 1][1]);
        if (wy == 0) row_to_lds(a.ep[0], pl + t.col - (uint32_t)a.nzi, &l_rho[nb][0][1]);
      } else {
        v.rho = *at(a.ep[0], pl + t.col);
        if (t.ty == 0) v.rho_hx = *at(a.ep[0], pl + t.col - (uint32_t)a.nzi);
      }
      if (t.tz == 0) v.rho_hz = *at(a.ep[0], pl + t.col - 1u);
    }
  };
  // the values fetch kept in registers, and the queue centres
  auto stage = [&](int b, const VelStage<T>& v, T cxy, T cyz) {
    if constexpr (!DMA) { l_xx[b][r][q] = v.xx; l_zz[b][r][q] = v.zz; l_xz[b][r][q] = v.xz; }
    l_xy[b][r][q] = cxy; l_yz[b][r][q] = cyz;
    if (!DMA && t.xh) {
      l_xx[b][t.ty][q] = v.hxx0; l_xx[b][r + BX][q] = v.hxx1;
      l_xy[b][t.ty][q] = v.hxy0; l_xy[b][r + BX][q] = v.hxy1;
      l_xz[b][t.ty][q] = v.hxz0; l_xz[b][r + BX][q] = v.hxz1;
    }
    if (t.zh) { l_zz[b][r][t.colh] = v.hzz; l_xz[b][r][t.colh] = v.hxz; l_yz[b][r][t.colh] = v.hyz; }
    if (MED3) {
      if constexpr (!DMA) {
        l_rho[b][t.ty + 1][t.tz + 1] = v.rho;
        if (t.ty == 0) l_rho[b][0][t.tz + 1] = v.rho_hx;
      }
      if (t.tz == 0) l_rho[b][t.ty + 1][0] = v.rho_hz;
    }
  };
  // F_MED3: density of the column on the previous plane (y - 1)
  T rho_prev = MED3 ? *at(a.ep[0], (uint32_t)(t.y0 - 1) * ys + t.col) : T(0);

  // queues: sxy (D+y, q[j] = f[iy-3+j]), syy (D-y, f[iy-4+j]), syz (D+y)
  T qxy[8], qyy[8], qyz[8];
#pragma unroll
  for (int j = 0; j < 7; ++j) {
    qxy[j] = S.sxy((uint32_t)(t.y0 - 3 + j) * ys + t.col);
    qyy[j] = S.syy((uint32_t)(t.y0 - 4 + j) * ys + t.col);
    qyz[j] = S.syz((uint32_t)(t.y0 - 3 + j) * ys + t.col);
  }
  {
    VelStage<T> v0;
    fetch(v0, t.y0, 0);
    qxy[7] = v0.nxy; qyy[7] = v0.nyy; qyz[7] = v0.nyz;
    stage(0, v0, qxy[3], qyz[3]);
  }
  __syncthreads();
  int b = 0;
  for (int iy = t.y0; iy < t.y1; ++iy) {
    const bool more = iy + 1 < t.y1;
    VelStage<T> nx_;
    if (more) fetch(nx_, iy + 1, b ^ 1);  // buffer b ^ 1 is free: every wave passed the last barrier
    // F_STORE with F_Q16: the whole wave of an interior row takes the branch
    // (the block maxima are wave reductions) and only the interior lanes store
    if (QS ? t.xin : t.active) {
      const uint32_t i = (uint32_t)iy * ys + t.col;
      const T* cx = a.cx; const T* cy = a.cy; const T* cz = a.cz;
      // FP32: every per-cell operand is loaded before the first store (the
      // stores could alias them as far as the compiler knows, and a load
      // issued after a store waits for it); FP64 loads them where they are
      // used, which holds fewer registers (2 waves per SIMD)
      T o3[3], m3[3], ss3[3], al3[3], ar3[3], g3[3], d3[3];
      auto load_ss = [&](int f) { return *at(a.ssl[f], i) * a.ssfl + *at(a.ssr[f], i) * a.ssfr; };
#pragma unroll
      for (int f = 0; f < 3; ++f) {
        o3[f] = ld_nt(a.o[f], i);
        if (!MED3 && MODE != MODE_BORN_ADJ) m3[f] = ld_nt(a.m[f], i);
        if constexpr (HOIST) {
          if (SS) ss3[f] = load_ss(f);
          if (ACC) { al3[f] = *at(a.accl[f], i); ar3[f] = *at(a.accr[f], i); }
          if (F & F_GRAD) {  // BX_/BY_/BZ_ == VX/VY/VZ
            g3[f] = *at(a.grad[f], i);
            if constexpr (Q16) d3[f] = load_dq16(a.dq16[f], a.dqs[f], i);
            else d3[f] = ld_nt(a.drv[f], i);
          }
        }
      }
      if (MED3) {
        const T rc = l_rho[b][t.ty + 1][t.tz + 1];
        m3[0] = stag_buoy(a, rc, l_rho[b][t.ty][t.tz + 1]);
        m3[1] = stag_buoy(a, rc, rho_prev);
        m3[2] = stag_buoy(a, rc, l_rho[b][t.ty + 1][t.tz]);
      }
      const T sx = dmx<T, BX>(l_xx[b], r, q, cx) + dq(qxy, cy) + dpz<T, BX>(l_xz[b], r, q, cz);
      const T sy = dpx<T, BX>(l_xy[b], r, q, cx) + dq(qyy, cy) + dpz<T, BX>(l_yz[b], r, q, cz);
      const T sz = dpx<T, BX>(l_xz[b], r, q, cx) + dq(qyz, cy) + dmz<T, BX>(l_zz[b], r, q, cz);
      T w = T(0);
      if (!QS || t.active) w = damp_w(a, t.izo, t.ix, iy, t.wy0, t.wny, t.wnx);
      const T s3[3] = {sx, sy, sz};
      if (F & F_STORE) {
#pragma unroll
        for (int f = 0; f < 3; ++f) {
          if constexpr (QS) store_dq16(a.dq16[f], a.dqs[f], i, t.active ? s3[f] : T(0));
          else st_nt(a.dst[f], i, s3[f]);
        }
      }
#pragma unroll
      for (int f = 0; f < 3; ++f) {
        T u;
        if (MODE == MODE_FWD) u = w * o3[f] + m3[f] * s3[f];
        else if (MODE == MODE_ADJ) u = w * o3[f] - m3[f] * s3[f];
        else u = w * o3[f] - s3[f];
        if constexpr (!HOIST) {
          if (SS) ss3[f] = load_ss(f);
          if (ACC) { al3[f] = *at(a.accl[f], i); ar3[f] = *at(a.accr[f], i); }
          if (F & F_GRAD) { g3[f] = *at(a.grad[f], i); d3[f] = ld_nt(a.drv[f], i); }
        }
        if (SS) u += ss3[f];
        u *= w;
        if (QS && !t.active) continue;
        st_nt(a.o[f], i, u);
        if (MODE == MODE_BORN_ADJ_PM) st_nt(a.qn[f], i, u * m3[f]);
        if (ACC) {
          *at(a.accl[f], i) = al3[f] + u * a.accfl;
          *at(a.accr[f], i) = ar3[f] + u * a.accfr;
        }
        if (F & F_GRAD) *at(a.grad[f], i) = g3[f] + u * d3[f];
      }
    }
    if (MED3) rho_prev = l_rho[b][t.ty + 1][t.tz + 1];
#pragma unroll
    for (int j = 0; j < 7; ++j) { qxy[j] = qxy[j + 1]; qyy[j] = qyy[j + 1]; qyz[j] = qyz[j + 1]; }
    if (more) {
      qxy[7] = nx_.nxy; qyy[7] = nx_.nyy; qyz[7] = nx_.nyz;
      stage(b ^ 1, nx_, qxy[3], qyz[3]);
    }
    __syncthreads();
    b ^= 1;
  }
}

// ---------------------------------------------------------------------------
// Stress update: six stresses from vx, vy, vz (same pipeline).
// ---------------------------------------------------------------------------
// The stress stepper's LDS: double-buffered planes of vx, vy, vz and (F_MED3)
// shear-modulus planes with the low x row and low z column.
template <typename T, int BX, bool MED3>
struct StressLds {
  Plane<T, BX> vx[2], vy[2], vz[2];
  T mu[2][MED3 ? BX + 1 : 1][MED3 ? TZ + 1 : 1];
};

template <typename T>
struct StressStage {
  T hlo[3], hhi[3], hz[3];   // x-halo rows and z-halo column of vx, vy, vz
  T nx, ny, nz;              // new queue values (vx: plane+3, vy: +4, vz: +3)
  T mu, mu_hx, mu_hz, mu_hc; // F_MED3: shear modulus of the column, row x0 - 1, z0 - 1 and both
};

template <typename T, int BX, int MODE, int F>
__global__ void __launch_bounds__(TZ * BX) k_stress(StepArgs<T> a) {
  constexpr bool SS = (F & F_SS) != 0, ACC = (F & F_ACC) != 0, MED3 = (F & F_MED3) != 0;
  constexpr bool Q16 = (F & F_Q16) != 0, QS = Q16 && (F & F_STORE) != 0;
  static_assert(!Q16 || sizeof(T) == 4, "F_Q16 is an FP32 format");
  static_assert(!QS || (!SS && !ACC && (F & F_GRAD) == 0 && MODE == MODE_FWD), "F_Q16 stores: MODE_FWD recomputation");
  constexpr bool HOIST = sizeof(T) == 4;  // per-cell operands loaded ahead of the stores
  static_assert(!MED3 || MODE == MODE_FWD || MODE == MODE_ADJ || MODE == MODE_BORN_ADJ_PM,
                "F_MED3 needs MODE_FWD, MODE_ADJ or MODE_BORN_ADJ_PM");
  __shared__ StressLds<T, BX, MED3> L;  // one LDS object (see k_velocity)
  auto& l_vx = L.vx; auto& l_vy = L.vy; auto& l_vz = L.vz; auto& l_mu = L.mu;
  const TileCoord t = tile_coord<BX>(a.nbz, a.nbx, a.nby, a.ychunk, a.nz, a.nzi, a.zoff, a.nx, a.ny, a.remap, a.xgroup, a.chunks);
  if (t.empty) return;  // block-uniform, before any barrier
  const uint32_t ys = (uint32_t)a.ys;
  const VelVal<T, MODE> V{a}, Vlo{a}, Vhi{a}, Vzh{a};
  // FP32 rows that are plain array values go straight from HBM to LDS (see k_velocity)
  constexpr bool DMA = ELA_LDS_DMA && sizeof(T) == 4 && MODE != MODE_BORN_ADJ;
  auto src = [&](int f) -> const T* { return MODE == MODE_BORN_ADJ_PM ? a.q[f] : a.c[f]; };
  const int r = t.ty + FAT, q = t.tz + FAT;
  const int wy = __builtin_amdgcn_readfirstlane(t.ty), wr = wy + FAT;  // wave-uniform row

  // plane p's values: LDS rows of buffer nb directly (DMA), the rest into v
  auto fetch = [&](StressStage<T>& v, int p, int nb) {
    const uint32_t pl = (uint32_t)p * ys;
    v.nx = V.v(VX, pl + 3 * ys + t.col);
    v.ny = V.v(VY, pl + 4 * ys + t.col);
    v.nz = V.v(VZ, pl + 3 * ys + t.col);
    if (wy < FAT) {  // the x-halo loaders (t.xh) are whole waves
      if constexpr (DMA) {
        row_to_lds(src(VX), pl + t.cxl, &l_vx[nb][wy][FAT]); row_to_lds(src(VX), pl + t.cxh, &l_vx[nb][wr + BX][FAT]);
        row_to_lds(src(VY), pl + t.cxl, &l_vy[nb][wy][FAT]); row_to_lds(src(VY), pl + t.cxh, &l_vy[nb][wr + BX][FAT]);
        row_to_lds(src(VZ), pl + t.cxl, &l_vz[nb][wy][FAT]); row_to_lds(src(VZ), pl + t.cxh, &l_vz[nb][wr + BX][FAT]);
      } else {
#pragma unroll
        for (int f = 0; f < 3; ++f) { v.hlo[f] = Vlo.v(f, pl + t.cxl); v.hhi[f] = Vhi.v(f, pl + t.cxh); }
      }
    }
    if (t.zh) {
#pragma unroll
      for (int f = 0; f < 3; ++f) v.hz[f] = Vzh.v(f, pl + t.czh);
    }
  };
  // F_MED3: shear modulus of plane p around the column
  auto fetch_mu = [&](StressStage<T>& v, int p, int nb) {
    const uint32_t pl = (uint32_t)p * ys, nzi = (uint32_t)a.nzi;
    if constexpr (DMA) {
      row_to_lds(a.ep[2], pl + t.col, &l_mu[nb][wy + 1][1]);
      if (wy == 0) row_to_lds(a.ep[2], pl + t.col - nzi, &l_mu[nb][0][1]);
    } else {
      v.mu = *at(a.ep[2], pl + t.col);
      if (t.ty == 0) v.mu_hx = *at(a.ep[2], pl + t.col - nzi);
    }
    if (t.tz == 0) v.mu_hz = *at(a.ep[2], pl + t.col - 1u);
    if (t.ty == 0 && t.tz == 0) v.mu_hc = *at(a.ep[2], pl + t.col - nzi - 1u);
  };
  auto stage_mu = [&](int m, const StressStage<T>& v) {
    if constexpr (!DMA) {
      l_mu[m][t.ty + 1][t.tz + 1] = v.mu;
      if (t.ty == 0) l_mu[m][0][t.tz + 1] = v.mu_hx;
    }
    if (t.tz == 0) l_mu[m][t.ty + 1][0] = v.mu_hz;
    if (t.ty == 0 && t.tz == 0) l_mu[m][0][0] = v.mu_hc;
  };
  // the values fetch kept in registers, and the queue centres
  auto stage = [&](int b, const StressStage<T>& v, T cvx, T cvy, T cvz) {
    l_vx[b][r][q] = cvx; l_vy[b][r][q] = cvy; l_vz[b][r][q] = cvz;
    if (!DMA && t.xh) {
      l_vx[b][t.ty][q] = v.hlo[0]; l_vx[b][r + BX][q] = v.hhi[0];
      l_vy[b][t.ty][q] = v.hlo[1]; l_vy[b][r + BX][q] = v.hhi[1];
      l_vz[b][t.ty][q] = v.hlo[2]; l_vz[b][r + BX][q] = v.hhi[2];
    }
    if (t.zh) { l_vx[b][r][t.colh] = v.hz[0]; l_vy[b][r][t.colh] = v.hz[1]; l_vz[b][r][t.colh] = v.hz[2]; }
  };

  // queues: vx (D-y, q[j] = f[iy-4+j]), vy (D+y, f[iy-3+j]), vz (D-y)
  T qx[8], qy[8], qz[8];
#pragma unroll
  for (int j = 0; j < 7; ++j) {
    qx[j] = V.v(VX, (uint32_t)(t.y0 - 4 + j) * ys + t.col);
    qy[j] = V.v(VY, (uint32_t)(t.y0 - 3 + j) * ys + t.col);
    qz[j] = V.v(VZ, (uint32_t)(t.y0 - 4 + j) * ys + t.col);
  }
  // F_MED3: the previous plane's terms carried in registers -- Sx(mu) of the
  // column, mu of the column and of its z - 1 neighbour -- which are the
  // current plane's values one iteration later (same operands, same order)
  T sx_prev = T(0), mu_prev = T(0), muz_prev = T(0);
  if (MED3) {
    const uint32_t pl = (uint32_t)(t.y0 - 1) * ys;
    mu_prev = *at(a.ep[2], pl + t.col);
    muz_prev = *at(a.ep[2], pl + t.col - 1u);
    sx_prev = T(0.5) * (mu_prev + *at(a.ep[2], pl + t.col - (uint32_t)a.nzi));
  }
  {
    StressStage<T> v0;
    fetch(v0, t.y0, 0);
    qx[7] = v0.nx; qy[7] = v0.ny; qz[7] = v0.nz;
    stage(0, v0, qx[4], qy[3], qz[4]);
    if (MED3) {
      fetch_mu(v0, t.y0, 0);
      stage_mu(0, v0);
    }
  }
  __syncthreads();
  int b = 0;
  for (int iy = t.y0; iy < t.y1; ++iy) {
    const bool more = iy + 1 < t.y1;
    StressStage<T> nx_;
    if (more) {  // buffer b ^ 1 is free: every wave passed the last barrier
      fetch(nx_, iy + 1, b ^ 1);
      if (MED3) fetch_mu(nx_, iy + 1, b ^ 1);
    }
    if (QS ? t.xin : t.active) {  // F_Q16 stores: see k_velocity
      const uint32_t i = (uint32_t)iy * ys + t.col;
      const T* cx = a.cx; const T* cy = a.cy; const T* cz = a.cz;
      // FP32: every per-cell operand is loaded before the first store (see k_velocity)
      T o6[6], ss6[6], al6[6], ar6[6], g5[5], d6[6];
      auto load_ss = [&](int f) { return *at(a.ssl[f], i) * a.ssfl + *at(a.ssr[f], i) * a.ssfr; };
      auto load_grad = [&]() {
#pragma unroll
        for (int k = 0; k < 6; ++k) {
          if constexpr (Q16) d6[k] = load_dq16(a.dq16[SXX + k], a.dqs[SXX + k], i);
          else d6[k] = ld_nt(a.drv[SXX + k], i);
        }
#pragma unroll
        for (int m = 0; m < 5; ++m) g5[m] = *at(a.grad[L2M + m], i);
      };
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        o6[k] = ld_nt(a.o[SXX + k], i);
        if constexpr (HOIST) {
          if (SS) ss6[k] = load_ss(SXX + k);
          if (ACC) { al6[k] = *at(a.accl[SXX + k], i); ar6[k] = *at(a.accr[SXX + k], i); }
        }
      }
      if constexpr (HOIST && (F & F_GRAD) != 0) load_grad();
      T l2m = 0, lam = 0, mxz = 0, mxy = 0, myz = 0;
      if (MED3) {
        const int rr = t.ty + 1, qq = t.tz + 1;
        const T mu_c = l_mu[b][rr][qq], mu_xm = l_mu[b][rr - 1][qq], mu_zm = l_mu[b][rr][qq - 1];
        const T mu_xzm = l_mu[b][rr - 1][qq - 1];
        const T lr = ld_nt(a.ep[1], i);
        normal_moduli(a.tdt, a.tdt64, lr, mu_c, l2m, lam);
        const T sx_c = T(0.5) * (mu_c + mu_xm), sx_zm = T(0.5) * (mu_zm + mu_xzm);
        const T sy_c = T(0.5) * (mu_c + mu_prev), sy_zm = T(0.5) * (mu_zm + muz_prev);
        mxz = (T(0.5) * (sx_c + sx_zm)) * a.tdt;
        mxy = (T(0.5) * (sx_c + sx_prev)) * a.tdt;
        myz = (T(0.5) * (sy_c + sy_zm)) * a.tdt;
      } else if (MODE != MODE_BORN_ADJ) {
        l2m = ld_nt(a.m[L2M], i); lam = ld_nt(a.m[LAM], i);
        mxz = ld_nt(a.m[MXZ], i); mxy = ld_nt(a.m[MXY], i); myz = ld_nt(a.m[MYZ], i);
      }
      const T exx = dpx<T, BX>(l_vx[b], r, q, cx);
      const T eyy = dq(qy, cy);
      const T ezz = dpz<T, BX>(l_vz[b], r, q, cz);
      const T gxy = dq(qx, cy) + dmx<T, BX>(l_vy[b], r, q, cx);
      const T gxz = dmz<T, BX>(l_vx[b], r, q, cz) + dmx<T, BX>(l_vz[b], r, q, cx);
      const T gyz = dq(qz, cy) + dmz<T, BX>(l_vy[b], r, q, cz);
      T w = T(0);
      if (!QS || t.active) w = damp_w(a, t.izo, t.ix, iy, t.wy0, t.wny, t.wnx);
      if (F & F_STORE) {
        if constexpr (QS) {
          const T e6[6] = {exx, eyy, ezz, gxz, gxy, gyz};
#pragma unroll
          for (int k = 0; k < 6; ++k) store_dq16(a.dq16[SXX + k], a.dqs[SXX + k], i, t.active ? e6[k] : T(0));
        } else {
          st_nt(a.dst[SXX], i, exx); st_nt(a.dst[SYY], i, eyy); st_nt(a.dst[SZZ], i, ezz);
          st_nt(a.dst[SXZ], i, gxz); st_nt(a.dst[SXY], i, gxy); st_nt(a.dst[SYZ], i, gyz);
        }
      }
      T u[6];
      if (MODE == MODE_BORN_ADJ || MODE == MODE_BORN_ADJ_PM) {
        u[0] = w * o6[0] - exx;
        u[1] = w * o6[1] - eyy;
        u[2] = w * o6[2] - ezz;
        u[3] = w * o6[3] - gxz;
        u[4] = w * o6[4] - gxy;
        u[5] = w * o6[5] - gyz;
      } else {
        const T sg = MODE == MODE_FWD ? T(1) : T(-1);
        u[0] = w * o6[0] + sg * (l2m * exx) + sg * (lam * (eyy + ezz));
        u[1] = w * o6[1] + sg * (l2m * eyy) + sg * (lam * (exx + ezz));
        u[2] = w * o6[2] + sg * (l2m * ezz) + sg * (lam * (exx + eyy));
        u[3] = w * o6[3] + sg * (mxz * gxz);
        u[4] = w * o6[4] + sg * (mxy * gxy);
        u[5] = w * o6[5] + sg * (myz * gyz);
      }
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        const int f = SXX + k;
        if constexpr (!HOIST) {
          if (SS) ss6[k] = load_ss(f);
          if (ACC) { al6[k] = *at(a.accl[f], i); ar6[k] = *at(a.accr[f], i); }
        }
        T v = u[k];
        if (SS) v += ss6[k];
        v *= w;
        u[k] = v;
        if (QS && !t.active) continue;
        st_nt(a.o[f], i, v);
        if (ACC) {
          *at(a.accl[f], i) = al6[k] + v * a.accfl;
          *at(a.accr[f], i) = ar6[k] + v * a.accfr;
        }
      }
      if (MODE == MODE_BORN_ADJ_PM) {
        st_nt(a.qn[SXX], i, u[0] * l2m + (u[1] + u[2]) * lam);
        st_nt(a.qn[SYY], i, u[1] * l2m + (u[0] + u[2]) * lam);
        st_nt(a.qn[SZZ], i, u[2] * l2m + (u[0] + u[1]) * lam);
        st_nt(a.qn[SXZ], i, u[3] * mxz);
        st_nt(a.qn[SXY], i, u[4] * mxy);
        st_nt(a.qn[SYZ], i, u[5] * myz);
      }
      if (F & F_GRAD) {
        // lamb2MuDtw multiplies exx, eyy, ezz in sxx, syy, szz; lambDtw the
        // other two normal strains; the shear moduli one shear strain each
        if constexpr (!HOIST) load_grad();
        const T dxx = d6[0], dyy = d6[1], dzz = d6[2];
        const T l2 = u[0] * dxx + u[1] * dyy + u[2] * dzz;
        *at(a.grad[L2M], i) = g5[0] + l2;
        *at(a.grad[LAM], i) = g5[1] + ((u[0] + u[1] + u[2]) * (dxx + dyy + dzz) - l2);
        *at(a.grad[MXZ], i) = g5[2] + u[3] * d6[3];
        *at(a.grad[MXY], i) = g5[3] + u[4] * d6[4];
        *at(a.grad[MYZ], i) = g5[4] + u[5] * d6[5];
      }
    }
#pragma unroll
    for (int j = 0; j < 7; ++j) { qx[j] = qx[j + 1]; qy[j] = qy[j + 1]; qz[j] = qz[j + 1]; }
    if (MED3) {
      // this plane's terms become the previous plane's (every thread, active or not)
      const int rr = t.ty + 1, qq = t.tz + 1;
      mu_prev = l_mu[b][rr][qq];
      muz_prev = l_mu[b][rr][qq - 1];
      sx_prev = T(0.5) * (mu_prev + l_mu[b][rr - 1][qq]);
    }
    if (more) {
      qx[7] = nx_.nx; qy[7] = nx_.ny; qz[7] = nx_.nz;
      stage(b ^ 1, nx_, qx[4], qy[3], qz[4]);
      if (MED3) stage_mu(b ^ 1, nx_);
    }
    __syncthreads();
    b ^= 1;
  }
}

// ---------------------------------------------------------------------------
// Sparse point kernels. One launch covers all 9 components.
// ---------------------------------------------------------------------------
template <typename T>
struct PointArgs {
  const int32_t* pos[9];   // internal-layout positions of the component's grid (-1: skip)
  int64_t n[9];
  int64_t off[10];         // prefix offsets over components
  T* field[9];             // wavefield component arrays
  T* tr[9];                // traces
  int64_t tr_stride_pt;    // element stride between points in tr (nts for [n][nts], 1 for [nts][n])
  int64_t tr_stride_t;     // element stride between time samples (1 for [n][nts], n for [nts][n])
  int its;
  T fa, fb;                // filter weights for its and its+1
  // optional coarse accumulation of the injected value (Born passes)
  T* accl[9];
  T* accr[9];
  T accfl, accfr;
  // optional gradient imaging of the injected value (adjoint-state gradient);
  // with dq16 set the derivative combinations are the F_Q16 int16 arrays
  const T* drv[9];
  const int16_t* dq16[9];
  const float* dqs[9];
  T* grad[8];
  // optional premultiplied level to keep consistent (MODE_BORN_ADJ_PM)
  T* q[9];
  const T* m[8];
  // damping (injection only)
  const T* taper;
  int min_pad, nz, nx, ny, nzi, zoff;
  const int* ytab;         // stacked windows: per plane {window y0, window ny, window nx, 0}, or null
  const int* its_dev;      // coarse sample index read from device memory (graph replays), or null
  // fine step k = *step_dev + step_off read from device memory (graph replays
  // of gradient steps), or null: the coarse sample is k / sub and the weights
  // are filt[i + fo] and filt[sub + 1 + i + fo] with i = k % sub (the host's
  // fa / fb of that step)
  const int* step_dev;
  const T* filt;
  int sub, fo, step_off;
};

// Coarse sample and time-filter weights of a point launch.
template <typename T>
__device__ __forceinline__ void point_time(const PointArgs<T>& a, int& its, T& fa, T& fb) {
  if (a.step_dev) {
    const int k = *a.step_dev + a.step_off;
    its = k / a.sub;
    const int i = k - its * a.sub + a.fo;
    fa = a.filt[i];
    fb = a.filt[a.sub + 1 + i];
  } else {
    its = a.its_dev ? *a.its_dev : a.its;
    fa = a.fa;
    fb = a.fb;
  }
}

template <typename T>
__device__ __forceinline__ int comp_of(const PointArgs<T>& a, int64_t g) {
  int c = 0;
#pragma unroll
  for (int k = 1; k < 9; ++k) c += (g >= a.off[k]) ? 1 : 0;
  return c;
}

template <typename T>
__device__ __forceinline__ T point_taper(const PointArgs<T>& a, int32_t p) {
  const int ys = a.nzi * a.nx;
  const int iy = p / ys;
  const int rem = p - iy * ys;
  const int ix = rem / a.nzi;
  const int iz = rem - ix * a.nzi - a.zoff;
  int wy0 = 0, wny = a.ny, wnx = a.nx;
  if (a.ytab) { wy0 = a.ytab[4 * iy]; wny = a.ytab[4 * iy + 1]; wnx = a.ytab[4 * iy + 2]; }
  int d = min(min(iz - FAT, ix - FAT), min(a.nz - iz - 1 - FAT, wnx - ix - 1 - FAT));
  d = min(d, min(iy - wy0 - FAT, wy0 + wny - iy - 1 - FAT));
  return d < a.min_pad ? a.taper[d] : T(1);
}

// field[pos] += w(pos) * (tr[its] fa + tr[its+1] fb)
// (ker_inject_source_* / ker_inject_data_*, kernelsElaGpu_3D.cu:11-52, with the
// sponge factor that the reference applies right after injection).
template <typename T, bool ACC, bool GRAD = false, bool PM = false>
__global__ void __launch_bounds__(256) k_inject(PointArgs<T> a) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.off[9]) return;
  const int c = comp_of(a, g);
  const int64_t k = g - a.off[c];
  const int32_t p = a.pos[c][k];
  if (p < 0) return;
  int its;
  T fa, fb;
  point_time(a, its, fa, fb);
  const T taper = point_taper(a, p);
  // the injected value of component cc at this point (the three normal
  // stresses share their point list, so index k is the same point for them)
  auto injected = [&](int cc) {
    const T* tr = a.tr[cc] + k * a.tr_stride_pt;
    return taper * (tr[(int64_t)its * a.tr_stride_t] * fa + tr[(int64_t)(its + 1) * a.tr_stride_t] * fb);
  };
  const T v = injected(c);
  a.field[c][p] += v;
  if (ACC && a.accl[c]) {
    a.accl[c][p] += v * a.accfl;
    a.accr[c][p] += v * a.accfr;
  }
  // Points are unique within a component's list, so every update below has
  // one writer; the normal stresses, coupled through lamb2MuDtw / lambDtw,
  // are updated together by the sxx thread, with the stepper's formulas.
  T vn[3] = {T(0), T(0), T(0)};
  if ((PM || GRAD) && c == SXX) { vn[0] = v; vn[1] = injected(SYY); vn[2] = injected(SZZ); }
  if (PM) {
    // keep q = M n
    if (c < 3) {
      a.q[c][p] += v * a.m[c][p];
    } else if (c == SXX) {
      const T l2 = a.m[L2M][p], la = a.m[LAM][p];
      a.q[SXX][p] += vn[0] * l2 + (vn[1] + vn[2]) * la;
      a.q[SYY][p] += vn[1] * l2 + (vn[0] + vn[2]) * la;
      a.q[SZZ][p] += vn[2] * l2 + (vn[0] + vn[1]) * la;
    } else if (c >= SXZ) {
      a.q[c][p] += v * a.m[MXZ + (c - SXZ)][p];
    }
  }
  if (GRAD) {
    // the stepper imaged the level before this injection; add the injected part
    auto drv = [&](int cc) -> T {
      return a.dq16[cc] ? T(load_dq16(a.dq16[cc], a.dqs[cc], (uint32_t)p)) : a.drv[cc][p];
    };
    if (c < 3) {
      a.grad[c][p] += v * drv(c);
    } else if (c == SXX) {
      const T dxx = drv(SXX), dyy = drv(SYY), dzz = drv(SZZ);
      const T l2 = vn[0] * dxx + vn[1] * dyy + vn[2] * dzz;
      a.grad[L2M][p] += l2;
      a.grad[LAM][p] += (vn[0] + vn[1] + vn[2]) * (dxx + dyy + dzz) - l2;
    } else if (c >= SXZ) {
      a.grad[MXZ + (c - SXZ)][p] += v * drv(c);
    }
  }
}

// tr[its] += field[pos] fa ; tr[its+1] += field[pos] fb
// (ker_record_interp_data_* / ker_record_source_*, kernelsElaGpu_3D.cu:61-104).
template <typename T>
__global__ void __launch_bounds__(256) k_record(PointArgs<T> a) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.off[9]) return;
  const int c = comp_of(a, g);
  const int64_t k = g - a.off[c];
  const int32_t p = a.pos[c][k];
  if (p < 0) return;
  const T v = a.field[c][p];
  int its;
  T fa, fb;
  point_time(a, its, fa, fb);
  T* tr = a.tr[c] + k * a.tr_stride_pt;
  tr[(int64_t)its * a.tr_stride_t] += v * fa;
  tr[(int64_t)(its + 1) * a.tr_stride_t] += v * fb;
}

// Out-of-place transpose of a [rows][cols] block (row stride ld_in) into
// [cols][rows] (row stride ld_out).
template <typename T>
__global__ void __launch_bounds__(256) k_transpose(const T* __restrict__ in, T* __restrict__ out, int64_t rows,
                                                   int64_t cols, int64_t ld_in, int64_t ld_out) {
  __shared__ T tile[32][33];
  const int64_t c0 = (int64_t)blockIdx.x * 32, r0 = (int64_t)blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int j = ty; j < 32; j += 8) {
    const int64_t rr = r0 + j, cc = c0 + tx;
    if (rr < rows && cc < cols) tile[j][tx] = in[rr * ld_in + cc];
  }
  __syncthreads();
  for (int j = ty; j < 32; j += 8) {
    const int64_t cc = c0 + j, rr = r0 + tx;
    if (rr < rows && cc < cols) out[cc * ld_out + rr] = tile[tx][j];
  }
}

// Born forward secondary sources for coarse sample `its`
// (imagingElaFwdGpu_3D, kernelsElaGpu_3D.cu:1232-1438). V* are the coarse
// vx, vy, vz slices at its-1 (o), its (c), its+1 (n); o/n null at the ends.
template <typename T>
struct ImagingArgs {
  const T* o[3];
  const T* c[3];
  const T* n[3];
  const T* dm[8];   // scaled perturbations (drhox drhoy drhoz dlame dmu dmuxz dmuxy dmuyz)
  double* dmo[8];   // adjoint: accumulated images (FP64: summed over every coarse sample)
  T* ss[9];         // forward: secondary sources out; adjoint: adjoint coarse slices in
  T cx[4], cy[4], cz[4];
  T dts_inv;
  int its, nts;
  int physics;      // 0: the yz term with the xz derivative pattern (reference); 1: the yz pattern
  int nz, nx, ny, nzi, zoff;
  int64_t ys;
  int ychunk, nbz, nbx, nby;
  int remap, xgroup;
  const int* chunks;
};

// The imaging kernels' LDS: double-buffered planes of the coarse vx, vy, vz
// slice V[its] (one LDS object, see k_velocity).
template <typename T, int BX>
struct ImagingLds {
  Plane<T, BX> vx[2], vy[2], vz[2];
};

// Secondary sources (forward) or images (adjoint) of coarse sample its. The
// strain rates of V[its] are those of the stress update (the same D+ / D-
// pattern), so the kernel runs the stress stepper's pipeline: planes of
// V[its] with x / z halos double-buffered in LDS (FP32 rows straight from
// HBM), y derivatives from 8-deep register queues, the next plane's loads
// issued before the current plane is computed, one barrier per plane, 32-bit
// offsets and unpredicated neighbour loads (every library array carries the
// guard that keeps them inside its allocation). Per interior cell it reads
// V[its - 1] and V[its + 1] (not at the ends of the record) and the 8
// perturbations (forward, writing 9 secondary sources) or the 9 adjoint
// slices and the 8 FP64 images (adjoint, rewriting the images).
template <typename T, int BX, bool ADJ>
__global__ void __launch_bounds__(TZ * BX) k_imaging(ImagingArgs<T> a) {
  __shared__ ImagingLds<T, BX> L;
  auto& l_vx = L.vx; auto& l_vy = L.vy; auto& l_vz = L.vz;
  const TileCoord t = tile_coord<BX>(a.nbz, a.nbx, a.nby, a.ychunk, a.nz, a.nzi, a.zoff, a.nx, a.ny, a.remap, a.xgroup, a.chunks);
  if (t.empty) return;  // block-uniform, before any barrier
  const uint32_t ys = (uint32_t)a.ys;
  constexpr bool DMA = ELA_LDS_DMA && sizeof(T) == 4;
  const int r = t.ty + FAT, q = t.tz + FAT;
  const int wy = __builtin_amdgcn_readfirstlane(t.ty), wr = wy + FAT;  // wave-uniform row
  const bool first = a.its == 0, last = a.its == a.nts - 1;            // kernel-uniform

  struct Stage {
    T hlo[3], hhi[3], hz[3];   // x-halo rows and z-halo column of vx, vy, vz
    T nx, ny, nz;              // new queue values (vx: plane + 3, vy: + 4, vz: + 3)
  };
  auto fetch = [&](Stage& v, int p, int nb) {
    const uint32_t pl = (uint32_t)p * ys;
    v.nx = *at(a.c[0], pl + 3 * ys + t.col);
    v.ny = *at(a.c[1], pl + 4 * ys + t.col);
    v.nz = *at(a.c[2], pl + 3 * ys + t.col);
    if (wy < FAT) {  // the x-halo loaders are whole waves
      if constexpr (DMA) {
        row_to_lds(a.c[0], pl + t.cxl, &l_vx[nb][wy][FAT]); row_to_lds(a.c[0], pl + t.cxh, &l_vx[nb][wr + BX][FAT]);
        row_to_lds(a.c[1], pl + t.cxl, &l_vy[nb][wy][FAT]); row_to_lds(a.c[1], pl + t.cxh, &l_vy[nb][wr + BX][FAT]);
        row_to_lds(a.c[2], pl + t.cxl, &l_vz[nb][wy][FAT]); row_to_lds(a.c[2], pl + t.cxh, &l_vz[nb][wr + BX][FAT]);
      } else {
#pragma unroll
        for (int f = 0; f < 3; ++f) { v.hlo[f] = *at(a.c[f], pl + t.cxl); v.hhi[f] = *at(a.c[f], pl + t.cxh); }
      }
    }
    if (t.zh) {
#pragma unroll
      for (int f = 0; f < 3; ++f) v.hz[f] = *at(a.c[f], pl + t.czh);
    }
  };
  auto stage = [&](int b, const Stage& v, T cvx, T cvy, T cvz) {
    l_vx[b][r][q] = cvx; l_vy[b][r][q] = cvy; l_vz[b][r][q] = cvz;
    if (!DMA && t.xh) {
      l_vx[b][t.ty][q] = v.hlo[0]; l_vx[b][r + BX][q] = v.hhi[0];
      l_vy[b][t.ty][q] = v.hlo[1]; l_vy[b][r + BX][q] = v.hhi[1];
      l_vz[b][t.ty][q] = v.hlo[2]; l_vz[b][r + BX][q] = v.hhi[2];
    }
    if (t.zh) { l_vx[b][r][t.colh] = v.hz[0]; l_vy[b][r][t.colh] = v.hz[1]; l_vz[b][r][t.colh] = v.hz[2]; }
  };

  // queues: vx (D-y, q[j] = f[iy-4+j]), vy (D+y, f[iy-3+j]), vz (D-y)
  T qx[8], qy[8], qz[8];
#pragma unroll
  for (int j = 0; j < 7; ++j) {
    qx[j] = *at(a.c[0], (uint32_t)(t.y0 - 4 + j) * ys + t.col);
    qy[j] = *at(a.c[1], (uint32_t)(t.y0 - 3 + j) * ys + t.col);
    qz[j] = *at(a.c[2], (uint32_t)(t.y0 - 4 + j) * ys + t.col);
  }
  {
    Stage v0;
    fetch(v0, t.y0, 0);
    qx[7] = v0.nx; qy[7] = v0.ny; qz[7] = v0.nz;
    stage(0, v0, qx[4], qy[3], qz[4]);
  }
  __syncthreads();
  int b = 0;
  for (int iy = t.y0; iy < t.y1; ++iy) {
    const bool more = iy + 1 < t.y1;
    Stage nx_;
    if (more) fetch(nx_, iy + 1, b ^ 1);  // buffer b ^ 1 is free: every wave passed the last barrier
    if (t.active) {
      const uint32_t i = (uint32_t)iy * ys + t.col;
      // every per-cell operand is loaded before the first store (a load
      // behind a store that might alias it waits for that store)
      T dv[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const T o = first ? T(0) : *at(a.o[k], i);
        const T n = last ? T(0) : *at(a.n[k], i);
        dv[k] = o - n;   // V[its-1] - V[its+1], zero outside the record
      }
      T dm[8], ss[9];
      double dmo[8];
      if constexpr (!ADJ) {
#pragma unroll
        for (int k = 0; k < 8; ++k) dm[k] = *at(a.dm[k], i);
      } else {
#pragma unroll
        for (int k = 0; k < 9; ++k) ss[k] = *at(a.ss[k], i);
#pragma unroll
        for (int k = 0; k < 8; ++k) dmo[k] = a.dmo[k][i];   // 64-bit indexing: FP64 offsets can pass 4 GiB
      }
      const T* cx = a.cx; const T* cy = a.cy; const T* cz = a.cz;
      const T exx = dpx<T, BX>(l_vx[b], r, q, cx);
      const T eyy = dq(qy, cy);
      const T ezz = dpz<T, BX>(l_vz[b], r, q, cz);
      const T gxy = dq(qx, cy) + dmx<T, BX>(l_vy[b], r, q, cx);
      const T gxz = dmz<T, BX>(l_vx[b], r, q, cz) + dmx<T, BX>(l_vz[b], r, q, cx);
      // yz term: the reference uses the xz derivative pattern
      // (kernelsElaGpu_3D.cu:1419-1431, :1622-1634); physics mode the yz one
      // of the stepper's syz update (D-y vz + D-z vy)
      const T gyz = a.physics ? dq(qz, cy) + dmz<T, BX>(l_vy[b], r, q, cz) : gxz;
      if constexpr (!ADJ) {
#pragma unroll
        for (int k = 0; k < 3; ++k) st_nt(a.ss[k], i, dm[k] * dv[k] * a.dts_inv);
        const T div = exx + eyy + ezz;
        st_nt(a.ss[SXX], i, dm[3] * div + T(2.0) * dm[4] * exx);
        st_nt(a.ss[SYY], i, dm[3] * div + T(2.0) * dm[4] * eyy);
        st_nt(a.ss[SZZ], i, dm[3] * div + T(2.0) * dm[4] * ezz);
        st_nt(a.ss[SXZ], i, dm[5] * gxz);
        st_nt(a.ss[SXY], i, dm[6] * gxy);
        st_nt(a.ss[SYZ], i, dm[7] * gyz);
      } else {
        // each coarse sample's term in T, the sum over samples in FP64
#pragma unroll
        for (int k = 0; k < 3; ++k) a.dmo[k][i] = dmo[k] + double(ss[k] * dv[k] * a.dts_inv);
        const T sxx = ss[SXX], syy = ss[SYY], szz = ss[SZZ];
        a.dmo[3][i] = dmo[3] + double((sxx + syy + szz) * (exx + eyy + ezz));
        a.dmo[4][i] = dmo[4] + double(T(2.0) * (exx * sxx + eyy * syy + ezz * szz));
        a.dmo[5][i] = dmo[5] + double(ss[SXZ] * gxz);
        a.dmo[6][i] = dmo[6] + double(ss[SXY] * gxy);
        a.dmo[7][i] = dmo[7] + double(ss[SYZ] * gyz);
      }
    }
#pragma unroll
    for (int j = 0; j < 7; ++j) { qx[j] = qx[j + 1]; qy[j] = qy[j + 1]; qz[j] = qz[j + 1]; }
    if (more) {
      qx[7] = nx_.nx; qy[7] = nx_.ny; qz[7] = nx_.nz;
      stage(b ^ 1, nx_, qx[4], qy[3], qz[4]);
    }
    __syncthreads();
    b ^= 1;
  }
}

// Reference-layout flat indices -> internal-layout offsets. Positions outside
// the FD interior are replaced by -1 (skipped by the point kernels) and
// reported through *bad.
__global__ void __launch_bounds__(256) k_convert_pos(const int64_t* __restrict__ in, int32_t* __restrict__ out,
                                                     int64_t n, int nz, int nx, int ny, int nzi, int zoff,
                                                     const int* ytab, int* bad) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t p = in[i];
  const int64_t ys = (int64_t)nz * nx;
  const int64_t iy = (p >= 0 && p < ys * ny) ? p / ys : -1;
  const int64_t r = p - iy * ys;
  const int64_t ix = r / nz, iz = r - ix * nz;
  bool ok = iy >= FAT && iy < ny - FAT && ix >= FAT && ix < nx - FAT && iz >= FAT && iz < nz - FAT;
  if (ok && ytab) {  // inside the interior of its window
    const int wy0 = ytab[4 * iy], wny = ytab[4 * iy + 1], wnx = ytab[4 * iy + 2];
    ok = wny > 0 && iy >= wy0 + FAT && iy < wy0 + wny - FAT && ix < wnx - FAT;
  }
  out[i] = ok ? (int32_t)(iy * nx * nzi + ix * nzi + iz + zoff) : -1;
  if (!ok) atomicOr(bad, 1);
}

// Device-side coarse-step counter of a replayed graph: *c += delta.
__global__ void k_count(int* c, int delta) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *c += delta;
}

template <typename T>
__global__ void __launch_bounds__(256) k_fill(T* p, int64_t n, T v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = i; k < n; k += stride) p[k] = v;
}

// ---------------------------------------------------------------------------
// Checkpoints of the forward state for the adjoint-state gradient: the 18
// arrays (two time levels x 9 fields) packed to the FD interior, either as
// they are or in bfloat16 scaled per array and per snapshot by a power of two
// 2^e >= max |x| (the scaling is exact; bfloat16 keeps 8 significant bits,
// rounded to nearest even).
// ---------------------------------------------------------------------------
// Unsigned 32-bit division by an invariant divisor d with a multiply-high,
// a subtraction and two shifts (Granlund and Montgomery, "Division by
// invariant integers using multiplication", PLDI 1994, section 4): with
// l = ceil(log2 d) and m = floor(2^32 (2^l - d) / d) + 1, t = mulhi(m, n),
// n / d = (t + ((n - t) >> min(l, 1))) >> max(l - 1, 0) for every 32-bit n.
// The host forms (m, l) once per grid (fast_div); the device avoids the
// 30-odd instructions of a general 32-bit division per element.
struct FastDiv {
  uint32_t m;
  int s1, s2;
};

__device__ __forceinline__ uint32_t fdiv(uint32_t n, FastDiv f) {
  const uint32_t t = __umulhi(f.m, n);
  return (t + ((n - t) >> f.s1)) >> f.s2;
}

template <typename T>
struct CkArgs {
  T* f[18];
  int nzI, nxI, nyI;     // interior extents (n - 2 FAT)
  int nzi, zoff;         // internal row pitch and z origin
  int64_t ys;            // nx * nzi
  int64_t nI;            // nzI * nxI * nyI
  FastDiv dz, dx;        // division by nzI and nxI
};

// Library offset of packed interior index q (32-bit: nI and every offset fit
// in 32 bits, the arrays being smaller than 4 GiB).
template <typename T>
__device__ __forceinline__ int64_t ck_offset(const CkArgs<T>& a, int64_t q) {
  const uint32_t q32 = (uint32_t)q, nz = (uint32_t)a.nzI, nx = (uint32_t)a.nxI;
  const uint32_t r = fdiv(q32, a.dz), iz = q32 - r * nz;
  const uint32_t iy = fdiv(r, a.dx), ix = r - iy * nx;
  return (int64_t)((iy + FAT) * (uint32_t)a.ys + (ix + FAT) * (uint32_t)a.nzi + (uint32_t)(a.zoff + FAT) + iz);
}

__device__ __forceinline__ uint16_t f32_to_bf16(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f32(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// Power-of-two scale >= amax (1 for an all-zero array).
__device__ __forceinline__ float pow2_scale(uint32_t amax_bits) {
  const float m = __uint_as_float(amax_bits);
  if (!(m > 0.f)) return 1.f;
  int e;
  const float fr = frexpf(m, &e);  // m = fr * 2^e, fr in [0.5, 1)
  (void)fr;
  return ldexpf(1.f, e);
}

// amax[f] = max |x| over the interior of array f (blockIdx.y = f), as the bit
// pattern of a non-negative float (ordered like unsigned integers).
template <typename T>
__global__ void __launch_bounds__(256) k_ck_amax(CkArgs<T> a, uint32_t* amax) {
  __shared__ float red[256];
  const int f = blockIdx.y;
  float m = 0.f;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < a.nI; q += (int64_t)gridDim.x * blockDim.x)
    m = fmaxf(m, fabsf((float)a.f[f][ck_offset(a, q)]));
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(&amax[f], __float_as_uint(red[0]));
}

// Pack: out[f][q] = x / scale_f as bfloat16 (BF16 = true) or x (BF16 = false).
template <typename T, bool BF16>
__global__ void __launch_bounds__(256) k_ck_pack(CkArgs<T> a, void* out, const uint32_t* amax, float* scale) {
  const int f = blockIdx.y;
  const float s = BF16 ? pow2_scale(amax[f]) : 1.f;
  const float inv = 1.f / s;   // exact: s is a power of two
  if (BF16 && blockIdx.x == 0 && threadIdx.x == 0) scale[f] = s;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < a.nI; q += (int64_t)gridDim.x * blockDim.x) {
    const T x = a.f[f][ck_offset(a, q)];
    if (BF16) static_cast<uint16_t*>(out)[f * a.nI + q] = f32_to_bf16((float)x * inv);
    else static_cast<T*>(out)[f * a.nI + q] = x;
  }
}

template <typename T, bool BF16>
__global__ void __launch_bounds__(256) k_ck_unpack(CkArgs<T> a, const void* in, const float* scale) {
  const int f = blockIdx.y;
  const float s = BF16 ? scale[f] : 1.f;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < a.nI; q += (int64_t)gridDim.x * blockDim.x) {
    T x;
    if (BF16) x = (T)(bf16_to_f32(static_cast<const uint16_t*>(in)[f * a.nI + q]) * s);
    else x = static_cast<const T*>(in)[f * a.nI + q];
    a.f[f][ck_offset(a, q)] = x;
  }
}

// The 8 scaled medium arrays from (rho, lambda, mu) on every cell of the
// padded grid with the host scaling's edge rules (the half-cell average
// halves the first sample and adds half the neighbour to the last one): the
// device twin of ela_scale_medium, same operations in the same order (FP32
// buoyancies and normal moduli formed in FP64 as in the steppers: buoy_h,
// normal_moduli).
template <typename T>
struct MedArgs {
  const T* ep[3];        // rho, lambda, mu (internal layout)
  T* m[8];               // the 8 scaled arrays (internal layout)
  int nz, nx, ny, nzi, zoff;
  int64_t ys;
  T tdt;
  double tdt64;
};

template <typename T>
__device__ __forceinline__ T half_cell(const T* f, int64_t i, int64_t s, int ia, int n) {
  return ia == 0 ? T(0.5) * f[i] : (ia < n - 1 ? T(0.5) * (f[i] + f[i - s]) : f[i] + T(0.5) * f[i - s]);
}
// the same in FP64 from the array's values (the FP32 buoyancies)
template <typename T>
__device__ __forceinline__ double half_cell64(const T* f, int64_t i, int64_t s, int ia, int n) {
  const double c = double(f[i]);
  return ia == 0 ? 0.5 * c : (ia < n - 1 ? 0.5 * (c + double(f[i - s])) : c + 0.5 * double(f[i - s]));
}

template <typename T>
__global__ void __launch_bounds__(256) k_scale_medium(MedArgs<T> a) {
  const int64_t n = (int64_t)a.nz * a.nx * a.ny;
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
    const int iz = (int)(c % a.nz);
    const int64_t rest = c / a.nz;
    const int ix = (int)(rest % a.nx), iy = (int)(rest / a.nx);
    const int64_t i = iy * a.ys + (int64_t)ix * a.nzi + a.zoff + iz;
    const int64_t sx = a.nzi, sy = a.ys;
    const T* rho = a.ep[0];
    const T* lam = a.ep[1];
    const T* mu = a.ep[2];
    if constexpr (sizeof(T) == 4) {
      a.m[BX_][i] = buoy_h<T>(a.tdt, a.tdt64, half_cell64(rho, i, sx, ix, a.nx));
      a.m[BY_][i] = buoy_h<T>(a.tdt, a.tdt64, half_cell64(rho, i, sy, iy, a.ny));
      a.m[BZ_][i] = buoy_h<T>(a.tdt, a.tdt64, half_cell64(rho, i, (int64_t)1, iz, a.nz));
    } else {
      a.m[BX_][i] = a.tdt / half_cell(rho, i, sx, ix, a.nx);
      a.m[BY_][i] = a.tdt / half_cell(rho, i, sy, iy, a.ny);
      a.m[BZ_][i] = a.tdt / half_cell(rho, i, (int64_t)1, iz, a.nz);
    }
    normal_moduli(a.tdt, a.tdt64, lam[i], mu[i], a.m[L2M][i], a.m[LAM][i]);
    // mu staggered along x, then along z (muxz) or y (muxy); along y, then z (muyz)
    const T tx = half_cell(mu, i, sx, ix, a.nx);
    const T tx_z = iz > 0 ? half_cell(mu, i - 1, sx, ix, a.nx) : T(0);
    const T tx_y = iy > 0 ? half_cell(mu, i - sy, sx, ix, a.nx) : T(0);
    const T ty = half_cell(mu, i, sy, iy, a.ny);
    const T ty_z = iz > 0 ? half_cell(mu, i - 1, sy, iy, a.ny) : T(0);
    auto outer = [](T v, T vm, int ia, int na) {
      return ia == 0 ? T(0.5) * v : (ia < na - 1 ? T(0.5) * (v + vm) : v + T(0.5) * vm);
    };
    a.m[MXZ][i] = outer(tx, tx_z, iz, a.nz) * a.tdt;
    a.m[MXY][i] = outer(tx, tx_y, iy, a.ny) * a.tdt;
    a.m[MYZ][i] = outer(ty, ty_z, iz, a.nz) * a.tdt;
  }
}

// Streaming reference for the roofline: R arrays of n4 16-byte vectors are
// read and the first W of them rewritten in place (non-temporal stores), the
// byte mix of a stencil kernel (reads of c, o and the medium, o rewritten)
// without neighbour accesses. Grid-stride loop over vectors. W = 0 reads
// only: the sum of all four lanes of the vector sum is compared with a value
// the (zero-filled) inputs never give, so every load stays live and nothing
// is written.
template <int R, int W>
__global__ void __launch_bounds__(256) k_stream(float4* const* arrays, int64_t n4) {
  float4* a[R];
#pragma unroll
  for (int k = 0; k < R; ++k) a[k] = arrays[k];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (int64_t)gridDim.x * blockDim.x) {
    float4 s = a[0][i];
#pragma unroll
    for (int k = 1; k < R; ++k) {
      const float4 v = a[k][i];
      s.x += v.x; s.y += v.y; s.z += v.z; s.w += v.w;
    }
    if constexpr (W == 0) {
      if (s.x + s.y + s.z + s.w == -1.2345e30f) a[0][i] = s;
    }
#pragma unroll
    for (int k = 0; k < W; ++k) {
      typedef float f4 __attribute__((ext_vector_type(4)));
      const f4 o = {s.x + (float)k, s.y, s.z, s.w};
      __builtin_nontemporal_store(o, reinterpret_cast<f4*>(a[k] + i));
    }
  }
}

// 16-bit checkpoints with one scale per block of CK_BLOCK consecutive packed
// values: q = rint(x / s) with s = max |x| / 32767 over the block (an
// all-zero block has s = 0), restored as x = q s. The absolute error is at
// most s / 2 (2^-16 of the block's largest value), against 2^-9 of each value
// for bfloat16. Layout of one checkpoint: int16 [18][nI], then float scales
// [18][ceil(nI / CK_BLOCK)] (36 nI bytes precede them, so they are 4-byte
// aligned). One wave quantises one block (the max is a wave reduction).
constexpr int CK_BLOCK = 64;

template <typename T>
__global__ void __launch_bounds__(256) k_ck_pack_s16(CkArgs<T> a, int16_t* out) {
  const int f = blockIdx.y;
  const int64_t nb = (a.nI + CK_BLOCK - 1) / CK_BLOCK;
  int16_t* q16 = out + f * a.nI;
  float* sc = reinterpret_cast<float*>(out + 18 * a.nI) + f * nb;
  const int lane = threadIdx.x & (CK_BLOCK - 1);
  const int64_t wstride = (int64_t)gridDim.x * (blockDim.x / CK_BLOCK);
  // b is uniform over a wave (blockDim.x is a multiple of 64), so every lane joins the reduction:
  // the maximum of each 32-lane half (DPP row moves and a swizzle), then of the two halves
  for (int64_t b = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / CK_BLOCK; b < nb; b += wstride) {
    const int64_t q = b * CK_BLOCK + lane;
    const bool ok = q < a.nI;
    const float x = ok ? (float)a.f[f][ck_offset(a, q)] : 0.f;
    const float h = half_wave_absmax(x);
    const float m = fmaxf(__int_as_float(__builtin_amdgcn_readlane(__float_as_int(h), 0)),
                          __int_as_float(__builtin_amdgcn_readlane(__float_as_int(h), 32)));
    const float inv = m > 0.f ? 32767.f / m : 0.f;
    if (ok) q16[q] = (int16_t)__float2int_rn(x * inv);
    if (lane == 0) sc[b] = m / 32767.f;
  }
}

template <typename T>
__global__ void __launch_bounds__(256) k_ck_unpack_s16(CkArgs<T> a, const int16_t* in) {
  const int f = blockIdx.y;
  const int64_t nb = (a.nI + CK_BLOCK - 1) / CK_BLOCK;
  const int16_t* q16 = in + f * a.nI;
  const float* sc = reinterpret_cast<const float*>(in + 18 * a.nI) + f * nb;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < a.nI; q += (int64_t)gridDim.x * blockDim.x)
    a.f[f][ck_offset(a, q)] = (T)((float)q16[q] * sc[q / CK_BLOCK]);
}

// Vectorised k_ck_pack_s16 / k_ck_unpack_s16 for FP32 grids whose interior
// z extent is a multiple of 8 (so are nI and every row start): a lane moves 8
// consecutive values -- 32 bytes of one z row (two 16-byte loads or stores)
// and 16 bytes of the packed int16 -- instead of one 4-byte and one 2-byte
// access, which kept too few bytes in flight per lane to reach the HBM rate.
// The 8 lanes of a 64-value block form its maximum with DPP moves inside
// their 8-lane group. Same blocks, scales and values as the scalar kernels.
__global__ void __launch_bounds__(256) k_ck_pack_s16_v8(CkArgs<float> a, int16_t* out) {
  const int f = blockIdx.y;
  const int64_t nb = (a.nI + CK_BLOCK - 1) / CK_BLOCK, n8 = a.nI / 8;
  int16_t* q16 = out + f * a.nI;
  float* sc = reinterpret_cast<float*>(out + 18 * a.nI) + f * nb;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  // g0 is uniform over the block, so every lane joins the reductions (lanes past the end hold zeros)
  for (int64_t g0 = (int64_t)blockIdx.x * blockDim.x; g0 < n8; g0 += stride) {
    const int64_t g = g0 + threadIdx.x;
    const bool ok = g < n8;
    float4 u = make_float4(0.f, 0.f, 0.f, 0.f), v = u;
    if (ok) {
      const float4* p = reinterpret_cast<const float4*>(a.f[f] + ck_offset(a, g * 8));
      u = p[0];
      v = p[1];
    }
    const float x[8] = {u.x, u.y, u.z, u.w, v.x, v.y, v.z, v.w};
    uint32_t mu = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) mu = max(mu, __float_as_uint(fabsf(x[k])));
    mu = dpp_umax<0xB1>(mu);   // quad_perm [1, 0, 3, 2]
    mu = dpp_umax<0x4E>(mu);   // quad_perm [2, 3, 0, 1]
    mu = dpp_umax<0x141>(mu);  // row_half_mirror: the 8-lane group's maximum
    const float m = __uint_as_float(mu);
    const float inv = m > 0.f ? 32767.f / m : 0.f;
    if (ok) {
      uint32_t w[4];
#pragma unroll
      for (int k = 0; k < 4; ++k)
        w[k] = (uint32_t)(uint16_t)(int16_t)__float2int_rn(x[2 * k] * inv) |
               ((uint32_t)(uint16_t)(int16_t)__float2int_rn(x[2 * k + 1] * inv) << 16);
      *reinterpret_cast<uint4*>(q16 + g * 8) = make_uint4(w[0], w[1], w[2], w[3]);
      if ((threadIdx.x & 7) == 0) sc[g / 8] = m / 32767.f;
    }
  }
}

__global__ void __launch_bounds__(256) k_ck_unpack_s16_v8(CkArgs<float> a, const int16_t* in) {
  const int f = blockIdx.y;
  const int64_t nb = (a.nI + CK_BLOCK - 1) / CK_BLOCK, n8 = a.nI / 8;
  const int16_t* q16 = in + f * a.nI;
  const float* sc = reinterpret_cast<const float*>(in + 18 * a.nI) + f * nb;
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < n8; g += (int64_t)gridDim.x * blockDim.x) {
    const uint4 w = *reinterpret_cast<const uint4*>(q16 + g * 8);
    const float s = sc[g / 8];
    auto lo = [&](uint32_t b) { return (float)(int16_t)(uint16_t)(b & 0xFFFFu) * s; };
    auto hi = [&](uint32_t b) { return (float)(int16_t)(uint16_t)(b >> 16) * s; };
    float4* p = reinterpret_cast<float4*>(a.f[f] + ck_offset(a, g * 8));
    p[0] = make_float4(lo(w.x), hi(w.x), lo(w.y), hi(w.y));
    p[1] = make_float4(lo(w.z), hi(w.z), lo(w.w), hi(w.w));
  }
}

// The gradient's FP32 partial sums folded into its FP64 sums over the FD
// interior only (the steppers and injections write nothing else): four
// consecutive z values per lane (interior z extent a multiple of 4, rows
// 16-byte aligned), array m = blockIdx.y of a.f (partial sums) and of acc
// (FP64 sums, stride np); the partial sums are cleared.
__global__ void __launch_bounds__(256) k_fold_v4(CkArgs<float> a, double* acc, int64_t np) {
  const int m = blockIdx.y;
  float* part = a.f[m];
  double* ac = acc + m * np;
  const int64_t n4 = a.nI / 4;
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < n4; g += (int64_t)gridDim.x * blockDim.x) {
    const int64_t o = ck_offset(a, g * 4);
    float4* pp = reinterpret_cast<float4*>(part + o);
    double2* d = reinterpret_cast<double2*>(ac + o);
    const float4 p = *pp;
    double2 d0 = d[0], d1 = d[1];
    d0.x += (double)p.x; d0.y += (double)p.y;
    d1.x += (double)p.z; d1.y += (double)p.w;
    d[0] = d0;
    d[1] = d1;
    *pp = make_float4(0.f, 0.f, 0.f, 0.f);
  }
}

// out[i] = (T) in[i] (the Born adjoint's FP64 image handed back in T).
template <typename T>
__global__ void __launch_bounds__(256) k_narrow(const double* in, T* out, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = T(in[i]);
}

// acc[i] += (double) part[i]; part[i] = 0 (FP64 time accumulation of the
// gradient from the per-segment partial sums).
template <typename T>
__global__ void __launch_bounds__(256) k_fold(T* part, double* acc, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    acc[i] += (double)part[i];
    part[i] = T(0);
  }
}

}  // namespace ela
