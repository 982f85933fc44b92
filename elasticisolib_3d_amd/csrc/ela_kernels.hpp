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
//   - block ids are remapped so that each XCD (blocks are dealt round-robin to
//     the 8 XCDs) owns a contiguous range of tiles: x/z neighbours that share
//     halo rows are resident on the same L2 at the same time.
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
       MODE_BORN_ADJ = 2   // n = w(w o - D (M c))      stepAdjGpuBorn_3D
};

template <typename T>
struct StepArgs {
  T* o[9];              // level n-1 in, level n+1 out (in place)
  const T* c[9];        // level n
  const T* m[8];        // scaled medium
  const T* ssl[9];      // secondary sources (left / right coarse slices)
  const T* ssr[9];
  T* accl[9];           // coarse-time accumulation targets
  T* accr[9];
  T ssfl, ssfr;         // filter weights for the secondary source
  T accfl, accfr;       // filter weights for the accumulation
  T cx[4], cy[4], cz[4];
  const T* taper;       // [min_pad]
  int min_pad;
  int nz, nx, ny;
  int64_t ys;           // nz * nx
  int ychunk;
  int nbz, nbx, nby;    // tile counts along z, x, y-chunks
};

__device__ __forceinline__ int xcd_remap(int bid, int nb) {
  const int per = nb / NXCD, rem = nb % NXCD;
  const int xcd = bid % NXCD, k = bid / NXCD;
  return xcd < rem ? xcd * (per + 1) + k : rem * (per + 1) + (xcd - rem) * per + k;
}

template <typename T>
__device__ __forceinline__ T damp_w(const StepArgs<T>& a, int iz, int ix, int iy) {
  int d = min(min(iz - FAT, ix - FAT), min(a.nz - iz - 1 - FAT, a.nx - ix - 1 - FAT));
  d = min(d, min(iy - FAT, a.ny - iy - 1 - FAT));
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

// Values that the derivative operators see. In MODE_BORN_ADJ they are the
// medium-multiplied fields (stepAdjGpuBorn_3D multiplies at every loaded
// position, kernelsElaGpu_3D.cu:864-1040).
template <typename T, int MODE>
struct StressVal {
  const StepArgs<T>& a;
  bool ok;
  __device__ __forceinline__ T sxx(int64_t i) const {
    if (!ok) return T(0);
    if (MODE == MODE_BORN_ADJ)
      return a.c[SXX][i] * a.m[L2M][i] + (a.c[SYY][i] + a.c[SZZ][i]) * a.m[LAM][i];
    return a.c[SXX][i];
  }
  __device__ __forceinline__ T syy(int64_t i) const {
    if (!ok) return T(0);
    if (MODE == MODE_BORN_ADJ)
      return a.c[SYY][i] * a.m[L2M][i] + (a.c[SXX][i] + a.c[SZZ][i]) * a.m[LAM][i];
    return a.c[SYY][i];
  }
  __device__ __forceinline__ T szz(int64_t i) const {
    if (!ok) return T(0);
    if (MODE == MODE_BORN_ADJ)
      return a.c[SZZ][i] * a.m[L2M][i] + (a.c[SXX][i] + a.c[SYY][i]) * a.m[LAM][i];
    return a.c[SZZ][i];
  }
  __device__ __forceinline__ T sxz(int64_t i) const {
    if (!ok) return T(0);
    return MODE == MODE_BORN_ADJ ? a.c[SXZ][i] * a.m[MXZ][i] : a.c[SXZ][i];
  }
  __device__ __forceinline__ T sxy(int64_t i) const {
    if (!ok) return T(0);
    return MODE == MODE_BORN_ADJ ? a.c[SXY][i] * a.m[MXY][i] : a.c[SXY][i];
  }
  __device__ __forceinline__ T syz(int64_t i) const {
    if (!ok) return T(0);
    return MODE == MODE_BORN_ADJ ? a.c[SYZ][i] * a.m[MYZ][i] : a.c[SYZ][i];
  }
};

template <typename T, int MODE>
struct VelVal {
  const StepArgs<T>& a;
  bool ok;
  __device__ __forceinline__ T v(int f, int64_t i) const {
    if (!ok) return T(0);
    return MODE == MODE_BORN_ADJ ? a.c[f][i] * a.m[f][i] : a.c[f][i];  // BX_/BY_/BZ_ == VX/VY/VZ
  }
};

struct TileCoord {
  int tz, ty, iz, ix, y0, y1, z0, x0;
};

template <int BX>
__device__ __forceinline__ TileCoord tile_coord(int nbz, int nbx, int nby, int ychunk, int ny) {
  const int nb = nbz * nbx * nby;
  const int lb = xcd_remap(blockIdx.x, nb);
  const int bz = lb % nbz;
  const int bx = (lb / nbz) % nbx;
  const int by = lb / (nbz * nbx);
  TileCoord t;
  t.tz = threadIdx.x;
  t.ty = threadIdx.y;
  t.z0 = FAT + bz * TZ;
  t.x0 = FAT + bx * BX;
  t.iz = t.z0 + t.tz;
  t.ix = t.x0 + t.ty;
  t.y0 = FAT + by * ychunk;
  t.y1 = min(t.y0 + ychunk, ny - FAT);
  return t;
}

// ---------------------------------------------------------------------------
// Velocity update: vx, vy, vz from the six stresses.
// ---------------------------------------------------------------------------
template <typename T, int BX, int MODE, bool SS, bool ACC>
__global__ void __launch_bounds__(TZ * BX) k_velocity(StepArgs<T> a) {
  __shared__ Plane<T, BX> l_xx, l_zz, l_xy, l_xz, l_yz;
  const TileCoord t = tile_coord<BX>(a.nbz, a.nbx, a.nby, a.ychunk, a.ny);
  const int nz = a.nz, nx = a.nx;
  const int64_t ys = a.ys;
  const bool okz = t.iz < nz, okx = t.ix < nx;
  const bool active = (t.iz < nz - FAT) && (t.ix < nx - FAT);
  const StressVal<T, MODE> S{a, okz && okx};
  // halo loaders: x-halo rows (ty < 4) and z-halo columns (tz < 8)
  const bool xh = t.ty < FAT;
  const int ixlo = t.x0 - FAT + t.ty, ixhi = t.x0 + BX + t.ty;
  const bool okxhi = ixhi < nx;
  const bool zh = t.tz < 2 * FAT;
  const int izh = t.tz < FAT ? t.z0 - FAT + t.tz : t.z0 + TZ + (t.tz - FAT);
  const int colh = t.tz < FAT ? t.tz : TZ + t.tz;
  const bool okzh = izh < nz;
  const StressVal<T, MODE> SXlo{a, okz && xh}, SXhi{a, okz && xh && okxhi};
  const StressVal<T, MODE> SZh{a, zh && okzh && okx};

  const int64_t col = (int64_t)t.ix * nz + t.iz;
  const int64_t colxlo = (int64_t)ixlo * nz + t.iz, colxhi = (int64_t)ixhi * nz + t.iz;
  const int64_t colzh = (int64_t)t.ix * nz + izh;

  T qxy[8], qyy[8], qyz[8];
#pragma unroll
  for (int j = 0; j < 7; ++j) {
    qxy[j] = S.sxy((int64_t)(t.y0 - 3 + j) * ys + col);
    qyy[j] = S.syy((int64_t)(t.y0 - 4 + j) * ys + col);
    qyz[j] = S.syz((int64_t)(t.y0 - 3 + j) * ys + col);
  }
  const int r = t.ty + FAT, q = t.tz + FAT;
  for (int iy = t.y0; iy < t.y1; ++iy) {
    const int64_t pl = (int64_t)iy * ys;
    const int64_t i = pl + col;
    qxy[7] = S.sxy(i + 4 * ys);
    qyy[7] = S.syy(i + 3 * ys);
    qyz[7] = S.syz(i + 4 * ys);
    const T cxx = S.sxx(i), czz = S.szz(i), cxz = S.sxz(i);
    T hxx_lo = 0, hxx_hi = 0, hxy_lo = 0, hxy_hi = 0, hxz_lo = 0, hxz_hi = 0;
    if (xh) {
      hxx_lo = SXlo.sxx(pl + colxlo); hxx_hi = SXhi.sxx(pl + colxhi);
      hxy_lo = SXlo.sxy(pl + colxlo); hxy_hi = SXhi.sxy(pl + colxhi);
      hxz_lo = SXlo.sxz(pl + colxlo); hxz_hi = SXhi.sxz(pl + colxhi);
    }
    T hzz = 0, hxzz = 0, hyzz = 0;
    if (zh) {
      hzz = SZh.szz(pl + colzh); hxzz = SZh.sxz(pl + colzh); hyzz = SZh.syz(pl + colzh);
    }
    __syncthreads();
    l_xx[r][q] = cxx; l_zz[r][q] = czz; l_xz[r][q] = cxz; l_xy[r][q] = qxy[3]; l_yz[r][q] = qyz[3];
    if (xh) {
      l_xx[t.ty][q] = hxx_lo; l_xx[r + BX][q] = hxx_hi;
      l_xy[t.ty][q] = hxy_lo; l_xy[r + BX][q] = hxy_hi;
      l_xz[t.ty][q] = hxz_lo; l_xz[r + BX][q] = hxz_hi;
    }
    if (zh) {
      l_zz[r][colh] = hzz; l_xz[r][colh] = hxzz; l_yz[r][colh] = hyzz;
    }
    __syncthreads();
    if (active) {
      const T* cx = a.cx; const T* cy = a.cy; const T* cz = a.cz;
      const T sx = dmx<T, BX>(l_xx, r, q, cx) + dq(qxy, cy) + dpz<T, BX>(l_xz, r, q, cz);
      const T sy = dpx<T, BX>(l_xy, r, q, cx) + dq(qyy, cy) + dpz<T, BX>(l_yz, r, q, cz);
      const T sz = dpx<T, BX>(l_xz, r, q, cx) + dq(qyz, cy) + dmz<T, BX>(l_zz, r, q, cz);
      const T w = damp_w(a, t.iz, t.ix, iy);
      const T s3[3] = {sx, sy, sz};
#pragma unroll
      for (int f = 0; f < 3; ++f) {
        T u;
        if (MODE == MODE_FWD) u = w * a.o[f][i] + a.m[f][i] * s3[f];
        else if (MODE == MODE_ADJ) u = w * a.o[f][i] - a.m[f][i] * s3[f];
        else u = w * a.o[f][i] - s3[f];
        if (SS) u += a.ssl[f][i] * a.ssfl + a.ssr[f][i] * a.ssfr;
        u *= w;
        a.o[f][i] = u;
        if (ACC) {
          a.accl[f][i] += u * a.accfl;
          a.accr[f][i] += u * a.accfr;
        }
      }
    }
#pragma unroll
    for (int j = 0; j < 7; ++j) { qxy[j] = qxy[j + 1]; qyy[j] = qyy[j + 1]; qyz[j] = qyz[j + 1]; }
  }
}

// ---------------------------------------------------------------------------
// Stress update: six stresses from vx, vy, vz.
// ---------------------------------------------------------------------------
template <typename T, int BX, int MODE, bool SS, bool ACC>
__global__ void __launch_bounds__(TZ * BX) k_stress(StepArgs<T> a) {
  __shared__ Plane<T, BX> l_vx, l_vy, l_vz;
  const TileCoord t = tile_coord<BX>(a.nbz, a.nbx, a.nby, a.ychunk, a.ny);
  const int nz = a.nz, nx = a.nx;
  const int64_t ys = a.ys;
  const bool okz = t.iz < nz, okx = t.ix < nx;
  const bool active = (t.iz < nz - FAT) && (t.ix < nx - FAT);
  const VelVal<T, MODE> V{a, okz && okx};
  const bool xh = t.ty < FAT;
  const int ixlo = t.x0 - FAT + t.ty, ixhi = t.x0 + BX + t.ty;
  const bool okxhi = ixhi < nx;
  const bool zh = t.tz < 2 * FAT;
  const int izh = t.tz < FAT ? t.z0 - FAT + t.tz : t.z0 + TZ + (t.tz - FAT);
  const int colh = t.tz < FAT ? t.tz : TZ + t.tz;
  const bool okzh = izh < nz;
  const VelVal<T, MODE> VXlo{a, okz && xh}, VXhi{a, okz && xh && okxhi};
  const VelVal<T, MODE> VZh{a, zh && okzh && okx};

  const int64_t col = (int64_t)t.ix * nz + t.iz;
  const int64_t colxlo = (int64_t)ixlo * nz + t.iz, colxhi = (int64_t)ixhi * nz + t.iz;
  const int64_t colzh = (int64_t)t.ix * nz + izh;

  // queues: vx (D-y), vy (D+y), vz (D-y)
  T qx[8], qy[8], qz[8];
#pragma unroll
  for (int j = 0; j < 7; ++j) {
    qx[j] = V.v(VX, (int64_t)(t.y0 - 4 + j) * ys + col);
    qy[j] = V.v(VY, (int64_t)(t.y0 - 3 + j) * ys + col);
    qz[j] = V.v(VZ, (int64_t)(t.y0 - 4 + j) * ys + col);
  }
  const int r = t.ty + FAT, q = t.tz + FAT;
  for (int iy = t.y0; iy < t.y1; ++iy) {
    const int64_t pl = (int64_t)iy * ys;
    const int64_t i = pl + col;
    qx[7] = V.v(VX, i + 3 * ys);
    qy[7] = V.v(VY, i + 4 * ys);
    qz[7] = V.v(VZ, i + 3 * ys);
    T hlo[3] = {0, 0, 0}, hhi[3] = {0, 0, 0}, hz[3] = {0, 0, 0};
    if (xh) {
#pragma unroll
      for (int f = 0; f < 3; ++f) { hlo[f] = VXlo.v(f, pl + colxlo); hhi[f] = VXhi.v(f, pl + colxhi); }
    }
    if (zh) {
#pragma unroll
      for (int f = 0; f < 3; ++f) hz[f] = VZh.v(f, pl + colzh);
    }
    __syncthreads();
    l_vx[r][q] = qx[4]; l_vy[r][q] = qy[3]; l_vz[r][q] = qz[4];
    if (xh) {
      l_vx[t.ty][q] = hlo[0]; l_vx[r + BX][q] = hhi[0];
      l_vy[t.ty][q] = hlo[1]; l_vy[r + BX][q] = hhi[1];
      l_vz[t.ty][q] = hlo[2]; l_vz[r + BX][q] = hhi[2];
    }
    if (zh) {
      l_vx[r][colh] = hz[0]; l_vy[r][colh] = hz[1]; l_vz[r][colh] = hz[2];
    }
    __syncthreads();
    if (active) {
      const T* cx = a.cx; const T* cy = a.cy; const T* cz = a.cz;
      const T exx = dpx<T, BX>(l_vx, r, q, cx);
      const T eyy = dq(qy, cy);
      const T ezz = dpz<T, BX>(l_vz, r, q, cz);
      const T gxy = dq(qx, cy) + dmx<T, BX>(l_vy, r, q, cx);
      const T gxz = dmz<T, BX>(l_vx, r, q, cz) + dmx<T, BX>(l_vz, r, q, cx);
      const T gyz = dq(qz, cy) + dmz<T, BX>(l_vy, r, q, cz);
      const T w = damp_w(a, t.iz, t.ix, iy);
      T u[6];
      if (MODE == MODE_BORN_ADJ) {
        u[0] = w * a.o[SXX][i] - exx;
        u[1] = w * a.o[SYY][i] - eyy;
        u[2] = w * a.o[SZZ][i] - ezz;
        u[3] = w * a.o[SXZ][i] - gxz;
        u[4] = w * a.o[SXY][i] - gxy;
        u[5] = w * a.o[SYZ][i] - gyz;
      } else {
        const T sg = MODE == MODE_FWD ? T(1) : T(-1);
        const T l2m = a.m[L2M][i], lam = a.m[LAM][i];
        u[0] = w * a.o[SXX][i] + sg * (l2m * exx) + sg * (lam * (eyy + ezz));
        u[1] = w * a.o[SYY][i] + sg * (l2m * eyy) + sg * (lam * (exx + ezz));
        u[2] = w * a.o[SZZ][i] + sg * (l2m * ezz) + sg * (lam * (exx + eyy));
        u[3] = w * a.o[SXZ][i] + sg * (a.m[MXZ][i] * gxz);
        u[4] = w * a.o[SXY][i] + sg * (a.m[MXY][i] * gxy);
        u[5] = w * a.o[SYZ][i] + sg * (a.m[MYZ][i] * gyz);
      }
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        const int f = SXX + k;
        T v = u[k];
        if (SS) v += a.ssl[f][i] * a.ssfl + a.ssr[f][i] * a.ssfr;
        v *= w;
        a.o[f][i] = v;
        if (ACC) {
          a.accl[f][i] += v * a.accfl;
          a.accr[f][i] += v * a.accfr;
        }
      }
    }
#pragma unroll
    for (int j = 0; j < 7; ++j) { qx[j] = qx[j + 1]; qy[j] = qy[j + 1]; qz[j] = qz[j + 1]; }
  }
}

// ---------------------------------------------------------------------------
// Sparse point kernels. One launch covers all 9 components.
// ---------------------------------------------------------------------------
template <typename T>
struct PointArgs {
  const int64_t* pos[9];   // positions of the component's grid
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
  // damping (injection only)
  const T* taper;
  int min_pad, nz, nx, ny;
};

template <typename T>
__device__ __forceinline__ int comp_of(const PointArgs<T>& a, int64_t g) {
  int c = 0;
#pragma unroll
  for (int k = 1; k < 9; ++k) c += (g >= a.off[k]) ? 1 : 0;
  return c;
}

template <typename T>
__device__ __forceinline__ T point_taper(const PointArgs<T>& a, int64_t p) {
  const int64_t ys = (int64_t)a.nz * a.nx;
  const int iy = (int)(p / ys);
  const int64_t rem = p - (int64_t)iy * ys;
  const int ix = (int)(rem / a.nz);
  const int iz = (int)(rem - (int64_t)ix * a.nz);
  int d = min(min(iz - FAT, ix - FAT), min(a.nz - iz - 1 - FAT, a.nx - ix - 1 - FAT));
  d = min(d, min(iy - FAT, a.ny - iy - 1 - FAT));
  return d < a.min_pad ? a.taper[d] : T(1);
}

// field[pos] += w(pos) * (tr[its] fa + tr[its+1] fb)
// (ker_inject_source_* / ker_inject_data_*, kernelsElaGpu_3D.cu:11-52, with the
// sponge factor that the reference applies right after injection).
template <typename T, bool ACC>
__global__ void __launch_bounds__(256) k_inject(PointArgs<T> a) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.off[9]) return;
  const int c = comp_of(a, g);
  const int64_t k = g - a.off[c];
  const int64_t p = a.pos[c][k];
  const T* tr = a.tr[c] + k * a.tr_stride_pt;
  const T s = tr[(int64_t)a.its * a.tr_stride_t] * a.fa + tr[(int64_t)(a.its + 1) * a.tr_stride_t] * a.fb;
  const T v = point_taper(a, p) * s;
  a.field[c][p] += v;
  if (ACC && a.accl[c]) {
    a.accl[c][p] += v * a.accfl;
    a.accr[c][p] += v * a.accfr;
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
  const T v = a.field[c][a.pos[c][k]];
  T* tr = a.tr[c] + k * a.tr_stride_pt;
  tr[(int64_t)a.its * a.tr_stride_t] += v * a.fa;
  tr[(int64_t)(a.its + 1) * a.tr_stride_t] += v * a.fb;
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
  T* dmo[8];        // adjoint: accumulated images
  T* ss[9];         // forward: secondary sources out; adjoint: adjoint coarse slices in
  T cx[4], cy[4], cz[4];
  T dts_inv;
  int its, nts;
  int nz, nx, ny;
  int64_t ys;
  int ychunk, nbz, nbx, nby;
};

template <typename T, int BX, bool ADJ>
__global__ void __launch_bounds__(TZ * BX) k_imaging(ImagingArgs<T> a) {
  __shared__ Plane<T, BX> l_vx, l_vy, l_vz;
  const TileCoord t = tile_coord<BX>(a.nbz, a.nbx, a.nby, a.ychunk, a.ny);
  const int nz = a.nz, nx = a.nx;
  const int64_t ys = a.ys;
  const bool ok = t.iz < nz && t.ix < nx;
  const bool active = (t.iz < nz - FAT) && (t.ix < nx - FAT);
  const bool xh = t.ty < FAT;
  const int ixlo = t.x0 - FAT + t.ty, ixhi = t.x0 + BX + t.ty;
  const bool okxlo = t.iz < nz && xh, okxhi = okxlo && ixhi < nx;
  const bool zh = t.tz < 2 * FAT;
  const int izh = t.tz < FAT ? t.z0 - FAT + t.tz : t.z0 + TZ + (t.tz - FAT);
  const int colh = t.tz < FAT ? t.tz : TZ + t.tz;
  const bool okzh = zh && izh < nz && t.ix < nx;
  const int64_t col = (int64_t)t.ix * nz + t.iz;
  const int64_t colxlo = (int64_t)ixlo * nz + t.iz, colxhi = (int64_t)ixhi * nz + t.iz;
  const int64_t colzh = (int64_t)t.ix * nz + izh;
  T qx[8], qy[8], qz[8];
#pragma unroll
  for (int j = 0; j < 7; ++j) {
    qx[j] = ld(a.c[0], (int64_t)(t.y0 - 4 + j) * ys + col, ok);
    qy[j] = ld(a.c[1], (int64_t)(t.y0 - 3 + j) * ys + col, ok);
    qz[j] = ld(a.c[2], (int64_t)(t.y0 - 4 + j) * ys + col, ok);
  }
  const int r = t.ty + FAT, q = t.tz + FAT;
  for (int iy = t.y0; iy < t.y1; ++iy) {
    const int64_t pl = (int64_t)iy * ys;
    const int64_t i = pl + col;
    qx[7] = ld(a.c[0], i + 3 * ys, ok);
    qy[7] = ld(a.c[1], i + 4 * ys, ok);
    qz[7] = ld(a.c[2], i + 3 * ys, ok);
    T hlo[3] = {0, 0, 0}, hhi[3] = {0, 0, 0}, hz[3] = {0, 0, 0};
    if (xh) {
#pragma unroll
      for (int f = 0; f < 3; ++f) { hlo[f] = ld(a.c[f], pl + colxlo, okxlo); hhi[f] = ld(a.c[f], pl + colxhi, okxhi); }
    }
    if (zh) {
#pragma unroll
      for (int f = 0; f < 3; ++f) hz[f] = ld(a.c[f], pl + colzh, okzh);
    }
    __syncthreads();
    l_vx[r][q] = qx[4]; l_vy[r][q] = qy[3]; l_vz[r][q] = qz[4];
    if (xh) {
      l_vx[t.ty][q] = hlo[0]; l_vx[r + BX][q] = hhi[0];
      l_vy[t.ty][q] = hlo[1]; l_vy[r + BX][q] = hhi[1];
      l_vz[t.ty][q] = hlo[2]; l_vz[r + BX][q] = hhi[2];
    }
    if (zh) { l_vx[r][colh] = hz[0]; l_vy[r][colh] = hz[1]; l_vz[r][colh] = hz[2]; }
    __syncthreads();
    if (active) {
      const T* cx = a.cx; const T* cy = a.cy; const T* cz = a.cz;
      const T exx = dpx<T, BX>(l_vx, r, q, cx);
      const T eyy = dq(qy, cy);
      const T ezz = dpz<T, BX>(l_vz, r, q, cz);
      const T gxy = dq(qx, cy) + dmx<T, BX>(l_vy, r, q, cx);
      const T gxz = dmz<T, BX>(l_vx, r, q, cz) + dmx<T, BX>(l_vz, r, q, cx);
      T dv[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        if (a.its == 0) dv[k] = -a.n[k][i];
        else if (a.its == a.nts - 1) dv[k] = a.o[k][i];
        else dv[k] = a.o[k][i] - a.n[k][i];
      }
      if (!ADJ) {
#pragma unroll
        for (int k = 0; k < 3; ++k) a.ss[k][i] = a.dm[k][i] * dv[k] * a.dts_inv;
        const T div = exx + eyy + ezz;
        a.ss[SXX][i] = a.dm[3][i] * div + T(2.0) * a.dm[4][i] * exx;
        a.ss[SYY][i] = a.dm[3][i] * div + T(2.0) * a.dm[4][i] * eyy;
        a.ss[SZZ][i] = a.dm[3][i] * div + T(2.0) * a.dm[4][i] * ezz;
        a.ss[SXZ][i] = a.dm[5][i] * gxz;
        a.ss[SXY][i] = a.dm[6][i] * gxy;
        // reference behaviour: the yz term uses the xz derivative pattern
        a.ss[SYZ][i] = a.dm[7][i] * gxz;
      } else {
#pragma unroll
        for (int k = 0; k < 3; ++k) a.dmo[k][i] += a.ss[k][i] * dv[k] * a.dts_inv;
        const T sxx = a.ss[SXX][i], syy = a.ss[SYY][i], szz = a.ss[SZZ][i];
        a.dmo[3][i] += (sxx + syy + szz) * (exx + eyy + ezz);
        a.dmo[4][i] += T(2.0) * (exx * sxx + eyy * syy + ezz * szz);
        a.dmo[5][i] += a.ss[SXZ][i] * gxz;
        a.dmo[6][i] += a.ss[SXY][i] * gxy;
        a.dmo[7][i] += a.ss[SYZ][i] * gxz;
      }
    }
#pragma unroll
    for (int j = 0; j < 7; ++j) { qx[j] = qx[j + 1]; qy[j] = qy[j + 1]; qz[j] = qz[j + 1]; }
  }
}

template <typename T>
__global__ void __launch_bounds__(256) k_fill(T* p, int64_t n, T v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = i; k < n; k += stride) p[k] = v;
}

}  // namespace ela
