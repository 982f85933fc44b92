// CPU restatement of the reference elastic isotropic 3-D propagator.
//
// TEST INFRASTRUCTURE ONLY. This file is the oracle that the parity tests,
// the package smoke check and the CPU-baseline leg of the benchmark use as the
// checker. Nothing in the product library links, loads or calls it.
//
// It restates, in plain C++ (templated on the scalar type, OpenMP over the
// slow axis), the algorithm of biondiettore/elasticIsoLib_3D:
//   - staggering (half-cell averages):     elastic_iso_float/src/stagger_3D.cpp:5-208
//   - medium staggering and scaling:       elastic_iso_float/src/fdParamElastic_3D.cpp:130-183
//   - FD coefficients, time filter, taper: elastic_iso_float/src/nonlinearPropElasticGpuFunctions_3D.cu:203-255
//                                          elastic_iso_float/src/BornElasticGpuFunctions_3D.cu:335-336 (dts_inv)
//   - stepper kernels:                     elastic_iso_float/src/kernelsElaGpu_3D.cu:110-1177
//   - damping:                             elastic_iso_float/src/kernelsElaGpu_3D.cu:1182-1226
//   - injection / recording:               elastic_iso_float/src/kernelsElaGpu_3D.cu:11-104
//   - Born imaging + coarse interpolation: elastic_iso_float/src/kernelsElaGpu_3D.cu:1232-1714
//   - nonlinear time loops:                elastic_iso_float/src/nonlinearPropElasticGpuFunctions_3D.cu:909-925, 1330-1345
//   - Born time loops:                     elastic_iso_float/src/BornElasticGpuFunctions_3D.cu:1268-1407, 2147-2281
//   - trilinear weights, unique points:    elastic_iso_float/src/spaceInterpGpu_3D.cpp:113-175, 214-279
//   - adjoint-state gradient (no reference counterpart): the transpose of the
//     linearised nl_forward, built from the pieces above (gradient()).
//
// Every loop is unfused and in the reference order (step, inject, damp,
// record/extract, swap), so that this file is a literal statement of the
// scheme rather than an optimised one. Layout: wavefields and medium arrays are
// [ny][nx][nz] (z fastest); traces are [nReg][nts] (time fastest), as in the
// reference's regular-grid trace buffers.
//
// Parity pins for each function are listed in tests/test_oracle.py.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <map>
#include <algorithm>
#include <omp.h>

namespace {

constexpr int FAT = 4;

// Component order of the model/data axis: vx vy vz sxx syy szz sxz sxy syz
// (elastic_iso_float/src/nonlinearPropElasticGpu_3D.cpp:297-333).
// Grid of each component: 0=centre 1=X 2=Y 3=Z 4=XZ 5=XY 6=YZ
// (python_float/Elastic_iso_float_3D.py:97-110).
constexpr int kCompGrid[9] = {1, 2, 3, 0, 0, 0, 4, 5, 6};

struct Grid {
  int64_t nz, nx, ny;
  int64_t nts, sub;  // sub = the doubled sub-sampling (host_sub)
  double dz, dx, dy, dts;
  int64_t minPad;
  double alphaCos;
};

// FD coefficients, literal values of
// elastic_iso_float/src/nonlinearPropElasticGpuFunctions_3D.cu:207-220.
// The y coefficients are divided by dx, not dy (reference behaviour).
const double kC[4] = {1.196289062541883, -0.079752604188901, 0.009570312506634,
                      -6.975446437140719e-04};

template <typename T>
struct Ctx {
  Grid g;
  int64_t nModel, ys, xs;  // total, y stride, x stride
  T cz[4], cx[4], cy[4];
  std::vector<T> filt;     // 2*(sub+1)
  std::vector<T> taper;    // minPad
  T dts_inv;
  explicit Ctx(const Grid& gg) : g(gg) {
    nModel = g.nz * g.nx * g.ny;
    xs = g.nz;
    ys = g.nz * g.nx;
    for (int k = 0; k < 4; ++k) {
      cz[k] = T(kC[k] / g.dz);
      cx[k] = T(kC[k] / g.dx);
      cy[k] = T(kC[k] / g.dx);
    }
    // Time filter: nonlinearPropElasticGpuFunctions_3D.cu:224-241
    const int64_t h = g.sub + 1;
    const int64_t ntw = (g.nts - 1) * g.sub + 1;
    filt.resize(2 * h);
    const double scale = 1.0 / std::sqrt(double(ntw) / double(g.nts));
    for (int64_t i = 0; i < h; ++i) {
      double a = 1.0 - double(i) / double(g.sub);
      filt[i] = T(a * scale);
      filt[h + i] = T((1.0 - a) * scale);
    }
    // Taper: nonlinearPropElasticGpuFunctions_3D.cu:248-255
    taper.resize(std::max<int64_t>(g.minPad, 1));
    for (int64_t d = 0; d < g.minPad; ++d) {
      double arg = M_PI / double(g.minPad) * double(g.minPad - d);
      taper[d] = T(g.alphaCos + (1.0 - g.alphaCos) * std::cos(arg));
    }
    // BornElasticGpuFunctions_3D.cu:335-336
    dts_inv = T(1.0 / (2.0 * g.dts));
  }
  int64_t h() const { return g.sub + 1; }
};

// D+ f[i] = sum_k c_k (f[i+k+1] - f[i-k]);  D- f[i] = sum_k c_k (f[i+k] - f[i-k-1])
// (kernelsElaGpu_3D.cu:326-453). Summation left to right as in the reference.
template <typename T>
inline T dplus(const T* f, int64_t i, int64_t s, const T* c) {
  return c[0] * (f[i + s] - f[i]) + c[1] * (f[i + 2 * s] - f[i - s]) +
         c[2] * (f[i + 3 * s] - f[i - 2 * s]) + c[3] * (f[i + 4 * s] - f[i - 3 * s]);
}
template <typename T>
inline T dminus(const T* f, int64_t i, int64_t s, const T* c) {
  return c[0] * (f[i] - f[i - s]) + c[1] * (f[i + s] - f[i - 2 * s]) +
         c[2] * (f[i + 2 * s] - f[i - 3 * s]) + c[3] * (f[i + 3 * s] - f[i - 4 * s]);
}

enum { VX = 0, VY, VZ, SXX, SYY, SZZ, SXZ, SXY, SYZ };
enum { BX = 0, BY, BZ, L2M, LAM, MXZ, MXY, MYZ };

template <typename T>
struct Fields {
  std::vector<T> f[9];
  explicit Fields(int64_t n) { for (auto& a : f) a.assign(n, T(0)); }
  void zero() { for (auto& a : f) std::fill(a.begin(), a.end(), T(0)); }
  void swap(Fields& o) { for (int i = 0; i < 9; ++i) f[i].swap(o.f[i]); }
};

// Interior loop helper: iy, ix, iz in [FAT, n-FAT).
template <typename T, typename F>
void for_interior(const Ctx<T>& C, F&& fn) {
  const Grid& g = C.g;
#pragma omp parallel for collapse(2) schedule(static)
  for (int64_t iy = FAT; iy < g.ny - FAT; ++iy)
    for (int64_t ix = FAT; ix < g.nx - FAT; ++ix) {
      const int64_t base = iy * C.ys + ix * C.xs;
      for (int64_t iz = FAT; iz < g.nz - FAT; ++iz) fn(base + iz);
    }
}

// Forward stepper (kernelsElaGpu_3D.cu:110-460) with sign=+1, nonlinear
// adjoint stepper (kernelsElaGpu_3D.cu:471-822) with sign=-1:
//   p0 <- p0 + sign * M D(p1)
template <typename T>
void step_md(const Ctx<T>& C, const T* const* m, Fields<T>& p0, const Fields<T>& p1, T sign) {
  const int64_t xs = C.xs, ys = C.ys;
  const T *cx = C.cx, *cy = C.cy, *cz = C.cz;
  const T *vx = p1.f[VX].data(), *vy = p1.f[VY].data(), *vz = p1.f[VZ].data();
  const T *sxx = p1.f[SXX].data(), *syy = p1.f[SYY].data(), *szz = p1.f[SZZ].data();
  const T *sxz = p1.f[SXZ].data(), *sxy = p1.f[SXY].data(), *syz = p1.f[SYZ].data();
  T* o[9];
  for (int i = 0; i < 9; ++i) o[i] = p0.f[i].data();
  for_interior(C, [&](int64_t i) {
    T dvx_dx = dplus(vx, i, xs, cx);
    T dvy_dy = dplus(vy, i, ys, cy);
    T dvz_dz = dplus(vz, i, 1, cz);
    o[VX][i] = o[VX][i] + sign * m[BX][i] *
               (dminus(sxx, i, xs, cx) + dplus(sxy, i, ys, cy) + dplus(sxz, i, 1, cz));
    o[VY][i] = o[VY][i] + sign * m[BY][i] *
               (dplus(sxy, i, xs, cx) + dminus(syy, i, ys, cy) + dplus(syz, i, 1, cz));
    o[VZ][i] = o[VZ][i] + sign * m[BZ][i] *
               (dplus(sxz, i, xs, cx) + dplus(syz, i, ys, cy) + dminus(szz, i, 1, cz));
    o[SXX][i] = o[SXX][i] + sign * m[L2M][i] * dvx_dx + sign * m[LAM][i] * (dvy_dy + dvz_dz);
    o[SYY][i] = o[SYY][i] + sign * m[L2M][i] * dvy_dy + sign * m[LAM][i] * (dvx_dx + dvz_dz);
    o[SZZ][i] = o[SZZ][i] + sign * m[L2M][i] * dvz_dz + sign * m[LAM][i] * (dvx_dx + dvy_dy);
    o[SXY][i] = o[SXY][i] + sign * m[MXY][i] * (dminus(vx, i, ys, cy) + dminus(vy, i, xs, cx));
    o[SXZ][i] = o[SXZ][i] + sign * m[MXZ][i] * (dminus(vx, i, 1, cz) + dminus(vz, i, xs, cx));
    o[SYZ][i] = o[SYZ][i] + sign * m[MYZ][i] * (dminus(vz, i, ys, cy) + dminus(vy, i, 1, cz));
  });
}

// Born adjoint stepper (kernelsElaGpu_3D.cu:825-1177): the exact transpose,
//   p0 <- p0 - D(M p1)   (medium applied before the derivative).
template <typename T>
void step_dm(const Ctx<T>& C, const T* const* m, Fields<T>& p0, const Fields<T>& p1) {
  const int64_t n = C.nModel;
  const int64_t xs = C.xs, ys = C.ys;
  const T *cx = C.cx, *cy = C.cy, *cz = C.cz;
  // Medium-multiplied copies of p1 (every cell, as the reference multiplies at
  // the neighbour positions it loads).
  std::vector<T> h[9];
  for (auto& a : h) a.resize(n);
  const T* const* c = nullptr;
  (void)c;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    h[VX][i] = p1.f[VX][i] * m[BX][i];
    h[VY][i] = p1.f[VY][i] * m[BY][i];
    h[VZ][i] = p1.f[VZ][i] * m[BZ][i];
    h[SXX][i] = p1.f[SXX][i] * m[L2M][i] + (p1.f[SYY][i] + p1.f[SZZ][i]) * m[LAM][i];
    h[SYY][i] = p1.f[SYY][i] * m[L2M][i] + (p1.f[SXX][i] + p1.f[SZZ][i]) * m[LAM][i];
    h[SZZ][i] = p1.f[SZZ][i] * m[L2M][i] + (p1.f[SXX][i] + p1.f[SYY][i]) * m[LAM][i];
    h[SXZ][i] = p1.f[SXZ][i] * m[MXZ][i];
    h[SXY][i] = p1.f[SXY][i] * m[MXY][i];
    h[SYZ][i] = p1.f[SYZ][i] * m[MYZ][i];
  }
  const T *vx = h[VX].data(), *vy = h[VY].data(), *vz = h[VZ].data();
  const T *sxx = h[SXX].data(), *syy = h[SYY].data(), *szz = h[SZZ].data();
  const T *sxz = h[SXZ].data(), *sxy = h[SXY].data(), *syz = h[SYZ].data();
  T* o[9];
  for (int i = 0; i < 9; ++i) o[i] = p0.f[i].data();
  for_interior(C, [&](int64_t i) {
    o[VX][i] = o[VX][i] - (dminus(sxx, i, xs, cx) + dplus(sxy, i, ys, cy) + dplus(sxz, i, 1, cz));
    o[VY][i] = o[VY][i] - (dplus(sxy, i, xs, cx) + dminus(syy, i, ys, cy) + dplus(syz, i, 1, cz));
    o[VZ][i] = o[VZ][i] - (dplus(sxz, i, xs, cx) + dplus(syz, i, ys, cy) + dminus(szz, i, 1, cz));
    o[SXX][i] = o[SXX][i] - dplus(vx, i, xs, cx);
    o[SYY][i] = o[SYY][i] - dplus(vy, i, ys, cy);
    o[SZZ][i] = o[SZZ][i] - dplus(vz, i, 1, cz);
    o[SXY][i] = o[SXY][i] - (dminus(vx, i, ys, cy) + dminus(vy, i, xs, cx));
    o[SXZ][i] = o[SXZ][i] - (dminus(vx, i, 1, cz) + dminus(vz, i, xs, cx));
    o[SYZ][i] = o[SYZ][i] - (dminus(vz, i, ys, cy) + dminus(vy, i, 1, cz));
  });
}

// dampCosineEdge_3D (kernelsElaGpu_3D.cu:1182-1226): both time levels.
template <typename T>
void damp(const Ctx<T>& C, Fields<T>& a, Fields<T>& b) {
  const Grid& g = C.g;
  if (g.minPad <= 0) return;
#pragma omp parallel for collapse(2) schedule(static)
  for (int64_t iy = FAT; iy < g.ny - FAT; ++iy)
    for (int64_t ix = FAT; ix < g.nx - FAT; ++ix)
      for (int64_t iz = FAT; iz < g.nz - FAT; ++iz) {
        int64_t d = std::min(std::min(iz - FAT, ix - FAT), std::min(g.nz - iz - 1 - FAT, g.nx - ix - 1 - FAT));
        d = std::min(d, std::min(iy - FAT, g.ny - iy - 1 - FAT));
        if (d < g.minPad) {
          const T w = C.taper[d];
          const int64_t i = iy * C.ys + ix * C.xs + iz;
          for (int k = 0; k < 9; ++k) {
            a.f[k][i] *= w;
            b.f[k][i] *= w;
          }
        }
      }
}

// Sparse point lists for the 7 grids (concatenated), traces for the 9
// components (concatenated, [nReg][nts] each).
template <typename T>
struct Points {
  const int64_t* pos[7];
  int64_t n[7];
  T* tr[9];
  Points(const int64_t* p, const int64_t* cnt, T* traces, int64_t nts) {
    int64_t off = 0;
    for (int gI = 0; gI < 7; ++gI) { pos[gI] = p + off; n[gI] = cnt[gI]; off += cnt[gI]; }
    int64_t toff = 0;
    for (int c = 0; c < 9; ++c) { tr[c] = traces + toff; toff += n[kCompGrid[c]] * nts; }
  }
};

// ker_inject_source_* (kernelsElaGpu_3D.cu:11-29) with filter index fi, and
// ker_inject_data_* (kernelsElaGpu_3D.cu:34-52) with fi = it2+1.
template <typename T>
void inject(const Ctx<T>& C, Fields<T>& p0, const Points<T>& P, int64_t its, int64_t fi) {
  const int64_t nts = C.g.nts, h = C.h();
  for (int c = 0; c < 9; ++c) {
    const int gI = kCompGrid[c];
    for (int64_t k = 0; k < P.n[gI]; ++k)
      p0.f[c][P.pos[gI][k]] += P.tr[c][nts * k + its] * C.filt[fi] + P.tr[c][nts * k + its + 1] * C.filt[h + fi];
  }
}

// ker_record_interp_data_* (kernelsElaGpu_3D.cu:84-104) and
// ker_record_source_* (kernelsElaGpu_3D.cu:61-79).
template <typename T>
void record(const Ctx<T>& C, const Fields<T>& p0, Points<T>& P, int64_t its, int64_t fi) {
  const int64_t nts = C.g.nts, h = C.h();
  for (int c = 0; c < 9; ++c) {
    const int gI = kCompGrid[c];
    for (int64_t k = 0; k < P.n[gI]; ++k) {
      const T v = p0.f[c][P.pos[gI][k]];
      P.tr[c][nts * k + its] += v * C.filt[fi];
      P.tr[c][nts * k + its + 1] += v * C.filt[h + fi];
    }
  }
}

// Nonlinear forward (nonlinearPropElasticGpuFunctions_3D.cu:876-975).
template <typename T>
void nl_forward(const Grid& g, const T* med, const T* src, const int64_t* spos, const int64_t* nsrc,
                T* rec, const int64_t* rpos, const int64_t* nrec) {
  Ctx<T> C(g);
  const T* m[8];
  for (int i = 0; i < 8; ++i) m[i] = med + i * C.nModel;
  Points<T> S(spos, nsrc, const_cast<T*>(src), g.nts);
  Points<T> R(rpos, nrec, rec, g.nts);
  Fields<T> p0(C.nModel), p1(C.nModel);
  for (int64_t its = 0; its < g.nts - 1; ++its)
    for (int64_t it2 = 1; it2 < g.sub + 1; ++it2) {
      step_md(C, m, p0, p1, T(1));
      inject(C, p0, S, its, it2 - 1);
      damp(C, p0, p1);
      record(C, p0, R, its, it2);
      p0.swap(p1);
    }
}

// Nonlinear adjoint (nonlinearPropElasticGpuFunctions_3D.cu:1298-1395).
template <typename T>
void nl_adjoint(const Grid& g, const T* med, T* src, const int64_t* spos, const int64_t* nsrc,
                const T* rec, const int64_t* rpos, const int64_t* nrec) {
  Ctx<T> C(g);
  const T* m[8];
  for (int i = 0; i < 8; ++i) m[i] = med + i * C.nModel;
  Points<T> S(spos, nsrc, src, g.nts);
  Points<T> R(rpos, nrec, const_cast<T*>(rec), g.nts);
  Fields<T> p0(C.nModel), p1(C.nModel);
  for (int64_t its = g.nts - 2; its > -1; --its)
    for (int64_t it2 = g.sub - 1; it2 > -1; --it2) {
      step_md(C, m, p0, p1, T(-1));
      inject(C, p0, R, its, it2 + 1);
      damp(C, p0, p1);
      record(C, p0, S, its, it2);
      p0.swap(p1);
    }
}

// Pass 1 of the Born operators: source wavefield with coarse-time
// accumulation of vx, vy, vz (BornElasticGpuFunctions_3D.cu:1268-1316,
// interpWavefieldVxVyVz_3D kernelsElaGpu_3D.cu:1646-1672).
// Returns V as [nts][3][nModel].
template <typename T>
std::vector<T> born_source_wavefield(const Ctx<T>& C, const T* const* m, const Points<T>& S) {
  const Grid& g = C.g;
  const int64_t n = C.nModel, h = C.h();
  std::vector<T> V(size_t(g.nts) * 3 * n, T(0));
  Fields<T> p0(n), p1(n);
  std::vector<T> L(3 * n, T(0)), Rr(3 * n, T(0));
  for (int64_t its = 0; its < g.nts - 1; ++its) {
    for (int64_t it2 = 1; it2 < g.sub + 1; ++it2) {
      step_md(C, m, p0, p1, T(1));
      inject(C, p0, S, its, it2 - 1);
      damp(C, p0, p1);
      for (int c = 0; c < 3; ++c) {
        T* l = L.data() + c * n;
        T* r = Rr.data() + c * n;
        const T* pf = p0.f[c].data();
        for_interior(C, [&](int64_t i) {
          l[i] += pf[i] * C.filt[it2];
          r[i] += pf[i] * C.filt[h + it2];
        });
      }
      p0.swap(p1);
    }
    std::copy(L.begin(), L.end(), V.begin() + its * 3 * n);
    L.swap(Rr);
    std::fill(Rr.begin(), Rr.end(), T(0));
  }
  std::copy(L.begin(), L.end(), V.begin() + (g.nts - 1) * 3 * n);
  return V;
}

enum { DRX = 0, DRY, DRZ, DLA, DMU, DMXZ, DMXY, DMYZ };

// The yz shear term of the Born secondary source and imaging. Reference mode
// (physics = false) keeps the reference's xz derivative pattern
// (kernelsElaGpu_3D.cu:1419-1431, :1622-1634); physics mode uses the yz
// pattern of the stepper's syz update (kernelsElaGpu_3D.cu:441-453), the true
// linearisation of the scheme. Nothing else differs between the two modes.
template <typename T>
inline T shear_yz(const T* cvx, const T* cvy, const T* cvz, int64_t i, int64_t xs, int64_t ys,
                  const T* cx, const T* cy, const T* cz, bool physics) {
  if (physics) return dminus(cvz, i, ys, cy) + dminus(cvy, i, 1, cz);
  return dminus(cvx, i, 1, cz) + dminus(cvz, i, xs, cx);
}

// imagingElaFwdGpu_3D (kernelsElaGpu_3D.cu:1232-1438). o/n may be null.
template <typename T>
void imaging_fwd(const Ctx<T>& C, const T* o, const T* c, const T* nn, int64_t its,
                 const T* const* dm, Fields<T>& ss, bool physics) {
  const int64_t n = C.nModel, xs = C.xs, ys = C.ys;
  const T *cx = C.cx, *cy = C.cy, *cz = C.cz;
  const T *cvx = c, *cvy = c + n, *cvz = c + 2 * n;
  const T dinv = C.dts_inv;
  const int64_t nts = C.g.nts;
  for_interior(C, [&](int64_t i) {
    T dvx_dx = dplus(cvx, i, xs, cx);
    T dvy_dy = dplus(cvy, i, ys, cy);
    T dvz_dz = dplus(cvz, i, 1, cz);
    for (int k = 0; k < 3; ++k) {
      T val;
      if (its == 0) val = dm[DRX + k][i] * (-nn[k * n + i]) * dinv;
      else if (its == nts - 1) val = dm[DRX + k][i] * (o[k * n + i]) * dinv;
      else val = dm[DRX + k][i] * (o[k * n + i] - nn[k * n + i]) * dinv;
      ss.f[VX + k][i] = val;
    }
    T div = dvx_dx + dvy_dy + dvz_dz;
    ss.f[SXX][i] = dm[DLA][i] * div + T(2.0) * dm[DMU][i] * dvx_dx;
    ss.f[SYY][i] = dm[DLA][i] * div + T(2.0) * dm[DMU][i] * dvy_dy;
    ss.f[SZZ][i] = dm[DLA][i] * div + T(2.0) * dm[DMU][i] * dvz_dz;
    ss.f[SXZ][i] = dm[DMXZ][i] * (dminus(cvx, i, 1, cz) + dminus(cvz, i, xs, cx));
    ss.f[SXY][i] = dm[DMXY][i] * (dminus(cvx, i, ys, cy) + dminus(cvy, i, xs, cx));
    // Reference quirk (reference mode): the yz secondary source uses the xz
    // derivative pattern (kernelsElaGpu_3D.cu:1419-1431).
    ss.f[SYZ][i] = dm[DMYZ][i] * shear_yz(cvx, cvy, cvz, i, xs, ys, cx, cy, cz, physics);
  });
}

// imagingElaAdjGpu_3D (kernelsElaGpu_3D.cu:1440-1641): accumulates into dm.
template <typename T>
void imaging_adj(const Ctx<T>& C, const T* o, const T* c, const T* nn, int64_t its,
                 T* const* dm, const Fields<T>& ss, bool physics) {
  const int64_t n = C.nModel, xs = C.xs, ys = C.ys;
  const T *cx = C.cx, *cy = C.cy, *cz = C.cz;
  const T *cvx = c, *cvy = c + n, *cvz = c + 2 * n;
  const T dinv = C.dts_inv;
  const int64_t nts = C.g.nts;
  for_interior(C, [&](int64_t i) {
    T dvx_dx = dplus(cvx, i, xs, cx);
    T dvy_dy = dplus(cvy, i, ys, cy);
    T dvz_dz = dplus(cvz, i, 1, cz);
    for (int k = 0; k < 3; ++k) {
      if (its == 0) dm[DRX + k][i] += ss.f[VX + k][i] * (-nn[k * n + i]) * dinv;
      else if (its == nts - 1) dm[DRX + k][i] += ss.f[VX + k][i] * (o[k * n + i]) * dinv;
      else dm[DRX + k][i] += ss.f[VX + k][i] * (o[k * n + i] - nn[k * n + i]) * dinv;
    }
    dm[DLA][i] += (ss.f[SXX][i] + ss.f[SYY][i] + ss.f[SZZ][i]) * (dvx_dx + dvy_dy + dvz_dz);
    dm[DMU][i] += T(2.0) * (dvx_dx * ss.f[SXX][i] + dvy_dy * ss.f[SYY][i] + dvz_dz * ss.f[SZZ][i]);
    dm[DMXZ][i] += ss.f[SXZ][i] * (dminus(cvx, i, 1, cz) + dminus(cvz, i, xs, cx));
    dm[DMXY][i] += ss.f[SXY][i] * (dminus(cvx, i, ys, cy) + dminus(cvy, i, xs, cx));
    // Same quirk as the forward imaging in reference mode (kernelsElaGpu_3D.cu:1622-1634).
    dm[DMYZ][i] += ss.f[SYZ][i] * shear_yz(cvx, cvy, cvz, i, xs, ys, cx, cy, cz, physics);
  });
}

// Born forward (BornElasticGpuFunctions_3D.cu:1226-1461).
template <typename T>
void born_forward(const Grid& g, const T* med, const T* src, const int64_t* spos, const int64_t* nsrc,
                  const T* dmod, T* rec, const int64_t* rpos, const int64_t* nrec, bool physics) {
  Ctx<T> C(g);
  const int64_t n = C.nModel, h = C.h();
  const T* m[8];
  const T* dm[8];
  for (int i = 0; i < 8; ++i) { m[i] = med + i * n; dm[i] = dmod + i * n; }
  Points<T> S(spos, nsrc, const_cast<T*>(src), g.nts);
  Points<T> R(rpos, nrec, rec, g.nts);
  std::vector<T> V = born_source_wavefield(C, m, S);
  auto Vs = [&](int64_t t) -> const T* { return (t < 0 || t >= g.nts) ? nullptr : V.data() + t * 3 * n; };
  Fields<T> p0(n), p1(n), ssL(n), ssR(n);
  imaging_fwd(C, Vs(-1), Vs(0), Vs(1), 0, dm, ssL, physics);
  for (int64_t its = 0; its < g.nts - 1; ++its) {
    imaging_fwd(C, Vs(its), Vs(its + 1), Vs(its + 2), its + 1, dm, ssR, physics);
    for (int64_t it2 = 1; it2 < g.sub + 1; ++it2) {
      step_md(C, m, p0, p1, T(1));
      // injectSecondarySource_3D (kernelsElaGpu_3D.cu:1675-1693), index it2-1
      for (int c = 0; c < 9; ++c) {
        T* pf = p0.f[c].data();
        const T* l = ssL.f[c].data();
        const T* r = ssR.f[c].data();
        const T fl = C.filt[it2 - 1], fr = C.filt[h + it2 - 1];
        for_interior(C, [&](int64_t i) { pf[i] += l[i] * fl + r[i] * fr; });
      }
      damp(C, p0, p1);
      record(C, p0, R, its, it2);
      p0.swap(p1);
    }
    ssL.swap(ssR);
    ssR.zero();
  }
}

// Born adjoint (BornElasticGpuFunctions_3D.cu:2103-2336). Accumulates into dmod.
template <typename T>
void born_adjoint(const Grid& g, const T* med, const T* src, const int64_t* spos, const int64_t* nsrc,
                  T* dmod, const T* rec, const int64_t* rpos, const int64_t* nrec, bool physics) {
  Ctx<T> C(g);
  const int64_t n = C.nModel, h = C.h();
  const T* m[8];
  T* dm[8];
  for (int i = 0; i < 8; ++i) { m[i] = med + i * n; dm[i] = dmod + i * n; }
  Points<T> S(spos, nsrc, const_cast<T*>(src), g.nts);
  Points<T> R(rpos, nrec, const_cast<T*>(rec), g.nts);
  std::vector<T> V = born_source_wavefield(C, m, S);
  auto Vs = [&](int64_t t) -> const T* { return (t < 0 || t >= g.nts) ? nullptr : V.data() + t * 3 * n; };
  Fields<T> p0(n), p1(n), ssL(n), ssR(n);
  for (int64_t its = g.nts - 2; its > -1; --its) {
    for (int64_t it2 = g.sub - 1; it2 > -1; --it2) {
      step_dm(C, m, p0, p1);
      inject(C, p0, R, its, it2 + 1);
      damp(C, p0, p1);
      // extractInterpAdjointWavefield_3D (kernelsElaGpu_3D.cu:1694-1714)
      for (int c = 0; c < 9; ++c) {
        const T* pf = p0.f[c].data();
        T* l = ssL.f[c].data();
        T* r = ssR.f[c].data();
        const T fl = C.filt[it2], fr = C.filt[h + it2];
        for_interior(C, [&](int64_t i) { l[i] += pf[i] * fl; r[i] += pf[i] * fr; });
      }
      p0.swap(p1);
    }
    imaging_adj(C, Vs(its), Vs(its + 1), Vs(its + 2), its + 1, dm, ssR, physics);
    ssR.swap(ssL);
    ssL.zero();
  }
  imaging_adj(C, Vs(-1), Vs(0), Vs(1), 0, dm, ssR, physics);
}

// ---- set-up helpers ------------------------------------------------------

// staggerX/Y/Z::adjoint (stagger_3D.cpp:47-70, 116-139, 185-208) and
// ::forward (stagger_3D.cpp:23-45, 92-114, 161-183). axis 0=z 1=x 2=y.
template <typename T>
void stagger(int axis, bool adjoint, int64_t nz, int64_t nx, int64_t ny, const T* in, T* out) {
  const int64_t s = axis == 0 ? 1 : axis == 1 ? nz : nz * nx;
  const int64_t na = axis == 0 ? nz : axis == 1 ? nx : ny;
  const int64_t n = nz * nx * ny;
  for (int64_t i = 0; i < n; ++i) out[i] = T(0);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    const int64_t a = (i / s) % na;
    if (!adjoint) {
      if (a < na - 1) out[i] += T(0.5) * (in[i] + in[i + s]);
      else out[i] += in[i];
    } else {
      if (a >= 1 && a < na - 1) out[i] += T(0.5) * (in[i] + in[i - s]);
      else if (a == 0) out[i] += T(0.5) * in[i];
      if (a == na - 1) out[i] += in[i] + T(0.5) * in[i - s];
    }
  }
}

// fdParamElastic_3D ctor (fdParamElastic_3D.cpp:130-183): rho, lambda, mu
// ([nModel] each) -> Bx By Bz L2M L Mxz Mxy Myz.
template <typename T>
void scale_medium(int64_t nz, int64_t nx, int64_t ny, double dtw, const T* rho, const T* lam,
                  const T* mu, T* out) {
  const int64_t n = nz * nx * ny;
  std::vector<T> t1(n), t2(n);
  T* o[8];
  for (int i = 0; i < 8; ++i) o[i] = out + i * n;
  stagger(1, true, nz, nx, ny, rho, o[0]);
  stagger(2, true, nz, nx, ny, rho, o[1]);
  stagger(0, true, nz, nx, ny, rho, o[2]);
  stagger(1, true, nz, nx, ny, mu, t1.data());
  stagger(0, true, nz, nx, ny, t1.data(), o[5]);
  stagger(2, true, nz, nx, ny, t1.data(), o[6]);
  stagger(2, true, nz, nx, ny, mu, t2.data());
  stagger(0, true, nz, nx, ny, t2.data(), o[7]);
  const T tdt = T(2.0 * dtw);
  for (int64_t i = 0; i < n; ++i) {
    o[0][i] = tdt / o[0][i];
    o[1][i] = tdt / o[1][i];
    o[2][i] = tdt / o[2][i];
    o[3][i] = tdt * (lam[i] + T(2.0) * mu[i]);
    o[4][i] = tdt * lam[i];
    o[5][i] = tdt * o[5][i];
    o[6][i] = tdt * o[6][i];
    o[7][i] = tdt * o[7][i];
  }
}

// Adjoint-state gradient of a data functional with respect to the 8 scaled
// medium arrays, for the nonlinear forward operator with fixed (already
// medium-scaled) source traces. No counterpart in the reference (its
// inversion images with the Born adjoint); this is the exact transpose of the
// linearisation of nl_forward above, in the same unfused order:
//   forward step j:  a' = W (a + M D b + s_j),   b' = W b,   record a'
//   dJ/dM           = sum_j  mu_j (.) D(b_j)   (per medium array, the
//                     derivative combination it multiplies in step_md)
//   mu_j            = the transposed recursion of born_adjoint's loop
//                     (step_dm, inject dJ/d(data), damp), whose p0 after
//                     damping is W lambda_j = mu_j
//   dJ/d(source traces) = mu_j extracted at the source points with the
//                     injection's filter weights (record, as nl_adjoint).
// `adj` holds dJ/d(receiver traces) [nReg][nts]; g8 and gsrc are accumulated.
template <typename T>
void gradient(const Grid& g, const T* med, const T* src, const int64_t* spos, const int64_t* nsrc, const T* adj,
              const int64_t* rpos, const int64_t* nrec, T* g8, T* gsrc) {
  Ctx<T> C(g);
  const int64_t n = C.nModel, xs = C.xs, ys = C.ys;
  const T* m[8];
  for (int i = 0; i < 8; ++i) m[i] = med + i * n;
  Points<T> S(spos, nsrc, const_cast<T*>(src), g.nts);
  Points<T> R(rpos, nrec, const_cast<T*>(adj), g.nts);
  Points<T> G(spos, nsrc, gsrc, g.nts);
  const int64_t nsteps = (g.nts - 1) * g.sub;
  std::vector<T> B(size_t(nsteps) * 9 * n);  // p1 before every forward step
  Fields<T> p0(n), p1(n);
  int64_t k = 0;
  for (int64_t its = 0; its < g.nts - 1; ++its)
    for (int64_t it2 = 1; it2 < g.sub + 1; ++it2, ++k) {
      for (int f = 0; f < 9; ++f) std::copy(p1.f[f].begin(), p1.f[f].end(), B.begin() + (k * 9 + f) * n);
      step_md(C, m, p0, p1, T(1));
      inject(C, p0, S, its, it2 - 1);
      damp(C, p0, p1);
      p0.swap(p1);
    }
  p0.zero();
  p1.zero();
  const T *cx = C.cx, *cy = C.cy, *cz = C.cz;
  for (int64_t its = g.nts - 2; its > -1; --its)
    for (int64_t it2 = g.sub - 1; it2 > -1; --it2) {
      step_dm(C, m, p0, p1);
      inject(C, p0, R, its, it2 + 1);
      damp(C, p0, p1);
      const T* b = B.data() + size_t(its * g.sub + it2) * 9 * n;
      const T *vx = b, *vy = b + n, *vz = b + 2 * n, *sxx = b + 3 * n, *syy = b + 4 * n, *szz = b + 5 * n;
      const T *sxz = b + 6 * n, *sxy = b + 7 * n, *syz = b + 8 * n;
      const T* u[9];
      for (int f = 0; f < 9; ++f) u[f] = p0.f[f].data();
      for_interior(C, [&](int64_t i) {
        const T sx = dminus(sxx, i, xs, cx) + dplus(sxy, i, ys, cy) + dplus(sxz, i, 1, cz);
        const T sy = dplus(sxy, i, xs, cx) + dminus(syy, i, ys, cy) + dplus(syz, i, 1, cz);
        const T sz = dplus(sxz, i, xs, cx) + dplus(syz, i, ys, cy) + dminus(szz, i, 1, cz);
        const T exx = dplus(vx, i, xs, cx), eyy = dplus(vy, i, ys, cy), ezz = dplus(vz, i, 1, cz);
        g8[BX * n + i] += u[VX][i] * sx;
        g8[BY * n + i] += u[VY][i] * sy;
        g8[BZ * n + i] += u[VZ][i] * sz;
        g8[L2M * n + i] += u[SXX][i] * exx + u[SYY][i] * eyy + u[SZZ][i] * ezz;
        g8[LAM * n + i] += u[SXX][i] * (eyy + ezz) + u[SYY][i] * (exx + ezz) + u[SZZ][i] * (exx + eyy);
        g8[MXY * n + i] += u[SXY][i] * (dminus(vx, i, ys, cy) + dminus(vy, i, xs, cx));
        g8[MXZ * n + i] += u[SXZ][i] * (dminus(vx, i, 1, cz) + dminus(vz, i, xs, cx));
        g8[MYZ * n + i] += u[SYZ][i] * (dminus(vz, i, ys, cy) + dminus(vy, i, 1, cz));
      });
      record(C, p0, G, its, it2);
      p0.swap(p1);
    }
}

}  // namespace

extern "C" {

struct orc_grid {
  int64_t nz, nx, ny, nts, sub;
  double dz, dx, dy, dts;
  int64_t minPad;
  double alphaCos;
};

static Grid to_grid(const orc_grid* p) {
  Grid g;
  g.nz = p->nz; g.nx = p->nx; g.ny = p->ny; g.nts = p->nts; g.sub = p->sub;
  g.dz = p->dz; g.dx = p->dx; g.dy = p->dy; g.dts = p->dts;
  g.minPad = p->minPad; g.alphaCos = p->alphaCos;
  return g;
}

#define ORC_INST(T, SFX)                                                                          \
  void orc_nl_forward_##SFX(const orc_grid* g, const T* med, const T* src, const int64_t* spos,  \
                            const int64_t* nsrc, T* rec, const int64_t* rpos, const int64_t* nrec) { \
    nl_forward<T>(to_grid(g), med, src, spos, nsrc, rec, rpos, nrec);                             \
  }                                                                                               \
  void orc_nl_adjoint_##SFX(const orc_grid* g, const T* med, T* src, const int64_t* spos,        \
                            const int64_t* nsrc, const T* rec, const int64_t* rpos,               \
                            const int64_t* nrec) {                                                \
    nl_adjoint<T>(to_grid(g), med, src, spos, nsrc, rec, rpos, nrec);                             \
  }                                                                                               \
  void orc_born_forward_##SFX(const orc_grid* g, const T* med, const T* src, const int64_t* spos, \
                              const int64_t* nsrc, const T* dmod, T* rec, const int64_t* rpos,    \
                              const int64_t* nrec, int physics) {                                 \
    born_forward<T>(to_grid(g), med, src, spos, nsrc, dmod, rec, rpos, nrec, physics != 0);       \
  }                                                                                               \
  void orc_born_adjoint_##SFX(const orc_grid* g, const T* med, const T* src, const int64_t* spos, \
                              const int64_t* nsrc, T* dmod, const T* rec, const int64_t* rpos,    \
                              const int64_t* nrec, int physics) {                                 \
    born_adjoint<T>(to_grid(g), med, src, spos, nsrc, dmod, rec, rpos, nrec, physics != 0);       \
  }                                                                                               \
  void orc_gradient_##SFX(const orc_grid* g, const T* med, const T* src, const int64_t* spos,       \
                          const int64_t* nsrc, const T* adj, const int64_t* rpos, const int64_t* nrec, \
                          T* g8, T* gsrc) {                                                       \
    gradient<T>(to_grid(g), med, src, spos, nsrc, adj, rpos, nrec, g8, gsrc);                     \
  }                                                                                               \
  void orc_stagger_##SFX(int axis, int adjoint, int64_t nz, int64_t nx, int64_t ny, const T* in,  \
                         T* out) {                                                                \
    stagger<T>(axis, adjoint != 0, nz, nx, ny, in, out);                                          \
  }                                                                                               \
  void orc_scale_medium_##SFX(int64_t nz, int64_t nx, int64_t ny, double dtw, const T* rho,       \
                              const T* lam, const T* mu, T* out) {                                \
    scale_medium<T>(nz, nx, ny, dtw, rho, lam, mu, out);                                          \
  }                                                                                               \
  void orc_step_##SFX(const orc_grid* g, const T* med, T* p0, const T* p1, int mode) {            \
    Grid gg = to_grid(g);                                                                         \
    Ctx<T> C(gg);                                                                                 \
    const T* m[8];                                                                                \
    for (int i = 0; i < 8; ++i) m[i] = med + i * C.nModel;                                        \
    Fields<T> a(C.nModel), b(C.nModel);                                                           \
    for (int k = 0; k < 9; ++k) {                                                                 \
      std::memcpy(a.f[k].data(), p0 + k * C.nModel, C.nModel * sizeof(T));                        \
      std::memcpy(b.f[k].data(), p1 + k * C.nModel, C.nModel * sizeof(T));                        \
    }                                                                                             \
    if (mode == 0) step_md(C, m, a, b, T(1));                                                     \
    else if (mode == 1) step_md(C, m, a, b, T(-1));                                               \
    else step_dm(C, m, a, b);                                                                     \
    for (int k = 0; k < 9; ++k) std::memcpy(p0 + k * C.nModel, a.f[k].data(), C.nModel * sizeof(T)); \
  }

ORC_INST(double, f64)
ORC_INST(float, f32)

// The Born imaging kernels alone, for pinning: V slices o, c, n (3 arrays
// each, o/n may be null at the ends as in the time loops), 8 perturbation
// arrays dm, 9 secondary-source arrays ss (forward: written; adjoint: read,
// images accumulated into dm). FP64 only.
void orc_imaging_fwd_f64(const orc_grid* g, const double* o, const double* c, const double* nn, int64_t its,
                         const double* dm, double* ss, int physics) {
  Ctx<double> C(to_grid(g));
  const double* d[8];
  for (int i = 0; i < 8; ++i) d[i] = dm + i * C.nModel;
  Fields<double> f(C.nModel);
  imaging_fwd(C, o, c, nn, its, d, f, physics != 0);
  for (int k = 0; k < 9; ++k) std::memcpy(ss + k * C.nModel, f.f[k].data(), C.nModel * sizeof(double));
}

void orc_imaging_adj_f64(const orc_grid* g, const double* o, const double* c, const double* nn, int64_t its,
                         double* dm, const double* ss, int physics) {
  Ctx<double> C(to_grid(g));
  double* d[8];
  for (int i = 0; i < 8; ++i) d[i] = dm + i * C.nModel;
  Fields<double> f(C.nModel);
  for (int k = 0; k < 9; ++k) std::memcpy(f.f[k].data(), ss + k * C.nModel, C.nModel * sizeof(double));
  imaging_adj(C, o, c, nn, its, d, f, physics != 0);
}

// Time filter and taper tables as the reference builds them (for pinning).
void orc_tables(const orc_grid* g, double* filt, double* taper) {
  Ctx<double> C(to_grid(g));
  for (size_t i = 0; i < C.filt.size(); ++i) filt[i] = C.filt[i];
  for (int64_t i = 0; i < g->minPad; ++i) taper[i] = C.taper[i];
}

int orc_num_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
