// Host-side set-up of the propagator: staggering and scaling of the elastic
// parameters (the work of fdParamElastic_3D's constructor,
// elastic_iso_float/src/fdParamElastic_3D.cpp:130-183, with the half-cell
// averages of elastic_iso_float/src/stagger_3D.cpp).
//
// Staggering rule along one axis (n samples, the adjoint half-cell average
// the reference applies to the density and shear modulus):
//   out[0] = in[0]/2,  out[i] = (in[i] + in[i-1])/2 for 0 < i < n-1,
//   out[n-1] = in[n-1] + in[n-2]/2.
// Scaled arrays: rho?Dtw = 2 dtw / rho_?, lamb2MuDtw = 2 dtw (lambda + 2 mu),
// lambDtw = 2 dtw lambda, mu??Dtw = 2 dtw mu_??.
#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <new>
#include <thread>
#include <vector>

#include "../../include/elastic_iso_3d.h"

namespace {

// Host worker threads: the hardware's, at most 16 and at most OMP_NUM_THREADS
// when that is set.
int host_threads() {
  int nt = static_cast<int>(std::max(1u, std::thread::hardware_concurrency()));
  nt = std::min(nt, 16);
  if (const char* e = std::getenv("OMP_NUM_THREADS")) {
    const int v = std::atoi(e);
    if (v > 0) nt = std::min(nt, v);
  }
  return nt;
}

// Runs an entry point's body with no exception crossing the C ABI: a failed
// host allocation is ELA_ERR_MEM, anything else ELA_ERR_ARG.
template <typename F>
int guarded(F f) {
  try {
    return f();
  } catch (const std::bad_alloc&) {
    return ELA_ERR_MEM;
  } catch (...) {
    return ELA_ERR_ARG;
  }
}

// f(begin, end) over [0, n) split into contiguous ranges, one per thread
// (serial below 64 Ki elements, each index standing for `work` elements).
// Every caller writes disjoint elements per index.
template <typename F>
void parallel_for(int64_t n, F f, int64_t work = 1) {
  const int64_t nt = host_threads();
  if (nt <= 1 || n < 2 || n * work < (int64_t(1) << 16)) {
    f(int64_t(0), n);
    return;
  }
  const int64_t chunk = (n + nt - 1) / nt;
  std::vector<std::thread> th;
  th.reserve(static_cast<size_t>(nt));
  int64_t b = 0;
  try {
    for (; b < n; b += chunk) th.emplace_back(f, b, std::min(n, b + chunk));
  } catch (...) {
    // a thread could not be started: the ranges left run on this thread
  }
  for (; b < n; b += chunk) f(b, std::min(n, b + chunk));
  for (auto& t : th) t.join();
}

// Adjoint half-cell average along an axis of stride s and length na on a
// [ny][nx][nz] array.
// The array is walked as [n / (s na)][na][s] blocks, so the axis index comes
// from the loop nest rather than a division per element. Threads take whole
// blocks, or the na rows of the single block when the axis is the slowest.
template <typename Row>
void along_axis(int64_t n, int64_t s, int64_t na, Row row) {
  const int64_t block = s * na, nblocks = n / block;
  if (nblocks >= 64) {
    parallel_for(nblocks, [&](int64_t b0, int64_t b1) {
      for (int64_t b = b0; b < b1; ++b)
        for (int64_t a = 0; a < na; ++a) row(b * block + a * s, a);
    }, block);
  } else {
    for (int64_t b = 0; b < nblocks; ++b)
      parallel_for(na, [&](int64_t a0, int64_t a1) {
        for (int64_t a = a0; a < a1; ++a) row(b * block + a * s, a);
      }, s);
  }
}

void half_cell_adjoint(const double* in, double* out, int64_t n, int64_t s, int64_t na) {
  along_axis(n, s, na, [=](int64_t off, int64_t a) {
    const double* r = in + off;
    double* w = out + off;
    if (a == 0) {
      for (int64_t j = 0; j < s; ++j) w[j] = 0.5 * r[j];
    } else if (a < na - 1) {
      for (int64_t j = 0; j < s; ++j) w[j] = 0.5 * (r[j] + r[j - s]);
    } else {
      for (int64_t j = 0; j < s; ++j) w[j] = r[j] + 0.5 * r[j - s];
    }
  });
}

// Forward half-cell average (the transpose of half_cell_adjoint):
//   out[i] = (in[i] + in[i+1])/2 for i < n-1,  out[n-1] = in[n-1].
void half_cell_forward(const double* in, double* out, int64_t n, int64_t s, int64_t na) {
  along_axis(n, s, na, [=](int64_t off, int64_t a) {
    const double* r = in + off;
    double* w = out + off;
    if (a < na - 1) {
      for (int64_t j = 0; j < s; ++j) w[j] = 0.5 * (r[j] + r[j + s]);
    } else {
      for (int64_t j = 0; j < s; ++j) w[j] = r[j];
    }
  });
}

bool dims_ok(int64_t nz, int64_t nx, int64_t ny) { return nz >= 2 && nx >= 2 && ny >= 2; }

}  // namespace

extern "C" int ela_scale_medium(int64_t nz, int64_t nx, int64_t ny, double dtw, const double* rho,
                                const double* lame, const double* mu, double* med8) {
  return guarded([&]() -> int {
    if (nz < 2 || nx < 2 || ny < 2 || !rho || !lame || !mu || !med8 || !(dtw > 0)) return ELA_ERR_ARG;
    const int64_t n = nz * nx * ny;
    const int64_t sz = 1, sx = nz, sy = nz * nx;
    double* Bx = med8;
    double* By = med8 + n;
    double* Bz = med8 + 2 * n;
    double* L2M = med8 + 3 * n;
    double* L = med8 + 4 * n;
    double* Mxz = med8 + 5 * n;
    double* Mxy = med8 + 6 * n;
    double* Myz = med8 + 7 * n;
    std::vector<double> t(n);
    half_cell_adjoint(rho, Bx, n, sx, nx);
    half_cell_adjoint(rho, By, n, sy, ny);
    half_cell_adjoint(rho, Bz, n, sz, nz);
    half_cell_adjoint(mu, t.data(), n, sx, nx);   // mu staggered along x
    half_cell_adjoint(t.data(), Mxz, n, sz, nz);  // ... then z
    half_cell_adjoint(t.data(), Mxy, n, sy, ny);  // ... then y
    half_cell_adjoint(mu, t.data(), n, sy, ny);   // mu staggered along y
    half_cell_adjoint(t.data(), Myz, n, sz, nz);  // ... then z
    const double tdt = 2.0 * dtw;
    parallel_for(n, [=](int64_t i0, int64_t i1) {
      for (int64_t i = i0; i < i1; ++i) {
        Bx[i] = tdt / Bx[i];
        By[i] = tdt / By[i];
        Bz[i] = tdt / Bz[i];
        L2M[i] = tdt * (lame[i] + 2.0 * mu[i]);
        L[i] = tdt * lame[i];
        Mxz[i] *= tdt;
        Mxy[i] *= tdt;
        Myz[i] *= tdt;
      }
    });
    return ELA_OK;
  });
}

// Born perturbations (BornElasticShotsGpu_3D.cpp:167-235): (drho, dlambda, dmu)
// -> drhox = rhoxDtw * Sx'(drho) (likewise y, z), dlame = 2 dtw dlambda,
// dmu = 2 dtw dmu, dmuxz = 2 dtw Sz'(Sx'(dmu)), dmuxy = 2 dtw Sy'(Sx'(dmu)),
// dmuyz = 2 dtw Sz'(Sy'(dmu)), with S' the half-cell average used for the
// medium (half_cell_adjoint).
extern "C" int ela_born_stagger(int64_t nz, int64_t nx, int64_t ny, double dtw, const double* med8,
                                const double* dm3, double* out8) {
  return guarded([&]() -> int {
    if (!dims_ok(nz, nx, ny) || !med8 || !dm3 || !out8 || !(dtw > 0)) return ELA_ERR_ARG;
    const int64_t n = nz * nx * ny, sz = 1, sx = nz, sy = nz * nx;
    const double* drho = dm3;
    const double* dlam = dm3 + n;
    const double* dmu = dm3 + 2 * n;
    const double tdt = 2.0 * dtw;
    std::vector<double> t(n);
    half_cell_adjoint(drho, out8, n, sx, nx);
    half_cell_adjoint(drho, out8 + n, n, sy, ny);
    half_cell_adjoint(drho, out8 + 2 * n, n, sz, nz);
    half_cell_adjoint(dmu, t.data(), n, sx, nx);
    half_cell_adjoint(t.data(), out8 + 5 * n, n, sz, nz);
    half_cell_adjoint(t.data(), out8 + 6 * n, n, sy, ny);
    half_cell_adjoint(dmu, t.data(), n, sy, ny);
    half_cell_adjoint(t.data(), out8 + 7 * n, n, sz, nz);
    parallel_for(n, [=](int64_t i0, int64_t i1) {
      for (int64_t i = i0; i < i1; ++i) {
        for (int k = 0; k < 3; ++k) out8[k * n + i] *= med8[k * n + i];
        out8[3 * n + i] = tdt * dlam[i];
        out8[4 * n + i] = tdt * dmu[i];
        for (int k = 5; k < 8; ++k) out8[k * n + i] *= tdt;
      }
    });
    return ELA_OK;
  });
}

// Adjoint of ela_born_stagger (BornElasticShotsGpu_3D.cpp:541-607): the 8
// images -> (drho, dlambda, dmu), overwritten.
extern "C" int ela_born_unstagger(int64_t nz, int64_t nx, int64_t ny, double dtw, const double* med8,
                                  const double* img8, double* out3) {
  return guarded([&]() -> int {
    if (!dims_ok(nz, nx, ny) || !med8 || !img8 || !out3 || !(dtw > 0)) return ELA_ERR_ARG;
    const int64_t n = nz * nx * ny, sz = 1, sx = nz, sy = nz * nx;
    const double tdt = 2.0 * dtw;
    std::vector<double> a(n), t(n), u(n);
    double* drho = out3;
    double* dlam = out3 + n;
    double* dmu = out3 + 2 * n;
    std::memset(drho, 0, n * sizeof(double));
    const int64_t st[3] = {sx, sy, sz}, na[3] = {nx, ny, nz};
    double* pa = a.data();
    const double* pt = t.data();
    const double* pu = u.data();
    for (int k = 0; k < 3; ++k) {
      parallel_for(n, [=](int64_t i0, int64_t i1) {
        for (int64_t i = i0; i < i1; ++i) pa[i] = img8[k * n + i] * med8[k * n + i];
      });
      half_cell_forward(a.data(), t.data(), n, st[k], na[k]);
      parallel_for(n, [=](int64_t i0, int64_t i1) {
        for (int64_t i = i0; i < i1; ++i) drho[i] += pt[i];
      });
    }
    parallel_for(n, [=](int64_t i0, int64_t i1) {
      for (int64_t i = i0; i < i1; ++i) {
        dlam[i] = tdt * img8[3 * n + i];
        dmu[i] = tdt * img8[4 * n + i];
      }
    });
    // (outer, inner) forward averages of muxz, muxy, muyz
    const int64_t so[3][2] = {{sx, sz}, {sx, sy}, {sy, sz}};
    const int64_t no[3][2] = {{nx, nz}, {nx, ny}, {ny, nz}};
    for (int k = 0; k < 3; ++k) {
      parallel_for(n, [=](int64_t i0, int64_t i1) {
        for (int64_t i = i0; i < i1; ++i) pa[i] = tdt * img8[(5 + k) * n + i];
      });
      half_cell_forward(a.data(), t.data(), n, so[k][1], no[k][1]);
      half_cell_forward(t.data(), u.data(), n, so[k][0], no[k][0]);
      parallel_for(n, [=](int64_t i0, int64_t i1) {
        for (int64_t i = i0; i < i1; ++i) dmu[i] += pu[i];
      });
    }
    return ELA_OK;
  });
}

// Gradient of a function of the 8 scaled medium arrays with respect to
// (rho, lambda, mu): the transposed Jacobian of ela_scale_medium applied to
// g8 = dJ/d(rhoxDtw ... muyzDtw). rho?Dtw = 2 dtw / S'(rho) gives
// d/drho = S(-(rho?Dtw)^2 / (2 dtw) g); lamb2MuDtw and lambDtw give
// 2 dtw (g_l2m + g_l) for lambda and 4 dtw g_l2m for mu; mu??Dtw give
// 2 dtw S_outer(S_inner(g)) for mu. out3 is overwritten.
extern "C" int ela_medium_gradient(int64_t nz, int64_t nx, int64_t ny, double dtw, const double* med8,
                                   const double* g8, double* out3) {
  return guarded([&]() -> int {
    if (!dims_ok(nz, nx, ny) || !med8 || !g8 || !out3 || !(dtw > 0)) return ELA_ERR_ARG;
    const int64_t n = nz * nx * ny, sz = 1, sx = nz, sy = nz * nx;
    const double tdt = 2.0 * dtw;
    std::vector<double> a(n), t(n), u(n);
    double* grho = out3;
    double* glam = out3 + n;
    double* gmu = out3 + 2 * n;
    std::memset(grho, 0, n * sizeof(double));
    const int64_t st[3] = {sx, sy, sz}, na[3] = {nx, ny, nz};
    double* pa = a.data();
    const double* pt = t.data();
    const double* pu = u.data();
    for (int k = 0; k < 3; ++k) {
      parallel_for(n, [=](int64_t i0, int64_t i1) {
        for (int64_t i = i0; i < i1; ++i) {
          const double b = med8[k * n + i];
          pa[i] = -b * b / tdt * g8[k * n + i];
        }
      });
      half_cell_forward(a.data(), t.data(), n, st[k], na[k]);
      parallel_for(n, [=](int64_t i0, int64_t i1) {
        for (int64_t i = i0; i < i1; ++i) grho[i] += pt[i];
      });
    }
    parallel_for(n, [=](int64_t i0, int64_t i1) {
      for (int64_t i = i0; i < i1; ++i) {
        glam[i] = tdt * (g8[3 * n + i] + g8[4 * n + i]);
        gmu[i] = 2.0 * tdt * g8[3 * n + i];
      }
    });
    const int64_t so[3][2] = {{sx, sz}, {sx, sy}, {sy, sz}};
    const int64_t no[3][2] = {{nx, nz}, {nx, ny}, {ny, nz}};
    for (int k = 0; k < 3; ++k) {
      parallel_for(n, [=](int64_t i0, int64_t i1) {
        for (int64_t i = i0; i < i1; ++i) pa[i] = tdt * g8[(5 + k) * n + i];
      });
      half_cell_forward(a.data(), t.data(), n, so[k][1], no[k][1]);
      half_cell_forward(t.data(), u.data(), n, so[k][0], no[k][0]);
      parallel_for(n, [=](int64_t i0, int64_t i1) {
        for (int64_t i = i0; i < i1; ++i) gmu[i] += pu[i];
      });
    }
    return ELA_OK;
  });
}

// ---------------------------------------------------------------------------
// Interpolation of acquisition devices onto one staggered grid: the linear and
// sinc methods of spaceInterpGpu_3D (elastic_iso_float/src/spaceInterpGpu_3D.cpp:
// constructor :13-79, bounds check :177-211, linear weights :214-279, dipole
// :281-346, sinc weights :351-480, unique points :113-138).
//
// The fractional position is formed in single precision, as the reference
// does (float wz = (z - oz) / dz; long long zReg = wz; wz = 1 - (wz - zReg)),
// so that grid-point indices agree with the reference for coordinates that
// sit on, or within rounding of, a grid node; weights are stored rounded to
// single precision as the reference stores them.
//
// Three corrections to the reference, which make its dipole and sinc paths
// usable: the sinc support check accepts a support that lies inside the
// stepped interior (the reference's comparison `i + h + 1 < n` is inverted
// and rejects every device, :371-382); the dipole pole's y weight divides by
// dy (the reference divides by dz, :287); the sinc dipole pole is negative as
// in the linear method (the reference leaves its sign positive, :470).
// ---------------------------------------------------------------------------
#include <cmath>
#include <unordered_map>

namespace {

constexpr double kPi = 3.14159265358979323846;

// sin(pi t) / (pi t), 1 at t = 0 (the reference's boost::math::sinc_pi(M_PI * t)).
double sinc_pi_t(double t) {
  if (t == 0.0) return 1.0;
  const double a = kPi * t;
  return std::sin(a) / a;
}

struct InterpBounds {
  int64_t lo[3], hi[3];  // allowed index range [lo, hi) per axis z, x, y
};

// Weights of one pole (sign +1 or -1) at (z, x, y) appended to (idx, w).
int pole_weights(const ela_interp_axes* ax, int method, int h, float z, float x, float y, double sign,
                 const InterpBounds& lin, const InterpBounds& snc, int64_t* idx, double* w) {
  const int64_t nz = ax->nz, nx = ax->nx;
  const float fz = (z - float(ax->oz)) / float(ax->dz);
  const float fx = (x - float(ax->ox)) / float(ax->dx);
  const float fy = (y - float(ax->oy)) / float(ax->dy);
  const int64_t zr = (int64_t)fz, xr = (int64_t)fx, yr = (int64_t)fy;
  if (method == ELA_INTERP_LINEAR) {
    if (yr < lin.lo[2] || yr + 1 >= lin.hi[2]) return ELA_ERR_ARG;
    if (xr < lin.lo[1] || xr + 1 >= lin.hi[1]) return ELA_ERR_ARG;
    if (zr < lin.lo[0] || zr + 1 >= lin.hi[0]) return ELA_ERR_ARG;
    const double az = 1.0f - (fz - float(zr)), ax_ = 1.0f - (fx - float(xr)), ay = 1.0f - (fy - float(yr));
    const int64_t b = yr * nz * nx + xr * nz + zr;
    const int64_t off[8] = {0, 1, nz, nz + 1, nz * nx, nz * nx + 1, nz * nx + nz, nz * nx + nz + 1};
    const double v[8] = {az * ax_ * ay,
                         (1.0 - az) * ax_ * ay,
                         az * (1.0 - ax_) * ay,
                         (1.0 - az) * (1.0 - ax_) * ay,
                         az * ax_ * (1.0 - ay),
                         (1.0 - az) * ax_ * (1.0 - ay),
                         az * (1.0 - ax_) * (1.0 - ay),
                         (1.0 - az) * (1.0 - ax_) * (1.0 - ay)};
    for (int k = 0; k < 8; ++k) {
      idx[k] = b + off[k];
      w[k] = double(float(sign * v[k]));
    }
    return ELA_OK;
  }
  // sinc: support zr-h+1 .. zr+h per axis, y slowest then x then z
  const int64_t r0[3] = {zr - h + 1, xr - h + 1, yr - h + 1};
  for (int a = 0; a < 3; ++a)
    if (r0[a] < snc.lo[a] || r0[a] + 2 * h > snc.hi[a]) return ELA_ERR_ARG;
  const int nf = 2 * h;
  std::vector<double> wz(nf), wx(nf), wy(nf);
  for (int k = 0; k < nf; ++k) {
    // argument (device - grid point) / d in single precision, as the reference forms it
    wz[k] = double((z - (float(r0[0] + k) * float(ax->dz) + float(ax->oz))) / float(ax->dz));
    wx[k] = double((x - (float(r0[1] + k) * float(ax->dx) + float(ax->ox))) / float(ax->dx));
    wy[k] = double((y - (float(r0[2] + k) * float(ax->dy) + float(ax->oy))) / float(ax->dy));
    wz[k] = sinc_pi_t(wz[k]);
    wx[k] = sinc_pi_t(wx[k]);
    wy[k] = sinc_pi_t(wy[k]);
  }
  int64_t e = 0;
  for (int iy = 0; iy < nf; ++iy)
    for (int ix = 0; ix < nf; ++ix)
      for (int iz = 0; iz < nf; ++iz, ++e) {
        idx[e] = nz * nx * (r0[2] + iy) + nz * (r0[1] + ix) + (r0[0] + iz);
        w[e] = double(float(sign * wz[iz] * wx[ix] * wy[iy]));
      }
  return ELA_OK;
}

}  // namespace

extern "C" int ela_interp_weights(const ela_interp_axes* ax, int64_t n_dev, const double* zc, const double* xc,
                                  const double* yc, int method, int half_length, int dipole, double z_shift,
                                  double x_shift, double y_shift, int64_t* nbr, double* weight,
                                  int64_t* unique_pos, int64_t* n_unique) {
  return guarded([&]() -> int {
    if (!ax || n_dev < 0 || (n_dev > 0 && (!zc || !xc || !yc || !nbr || !weight || !unique_pos || !n_unique)))
      return ELA_ERR_ARG;
    if (method != ELA_INTERP_LINEAR && method != ELA_INTERP_SINC) return ELA_ERR_ARG;
    if (method == ELA_INTERP_LINEAR && half_length != 1) return ELA_ERR_ARG;
    if (half_length < 1 || half_length > 32 || (dipole != 0 && dipole != 1)) return ELA_ERR_ARG;
    const int64_t nz = ax->nz, nx = ax->nx, ny = ax->ny, fat = ax->fat;
    // device bounds (checkOutOfBounds, spaceInterpGpu_3D.cpp:177-211)
    const float zMin = float(ax->oz + (fat + ax->z_pad_minus) * ax->dz);
    const float xMin = float(ax->ox + (fat + ax->x_pad_minus) * ax->dx);
    const float yMin = float(ax->oy + (fat + ax->y_pad) * ax->dy);
    const float zMax = zMin + float((nz - 2 * fat - ax->z_pad_minus - ax->z_pad_plus - 1) * ax->dz);
    const float xMax = xMin + float((nx - 2 * fat - ax->x_pad_minus - ax->x_pad_plus - 1) * ax->dx);
    const float yMax = yMin + float((ny - 2 * fat - 2 * ax->y_pad - 1) * ax->dy);
    const float tol = float(ax->tolerance);
    // linear: both neighbours inside the unpadded model; sinc: the support inside the stepped interior
    const InterpBounds lin{{fat + ax->z_pad_minus, fat + ax->x_pad_minus, fat + ax->y_pad},
                           {nz - fat - ax->z_pad_plus, nx - fat - ax->x_pad_plus, ny - fat - ax->y_pad}};
    const InterpBounds snc{{fat, fat, fat}, {nz - fat, nx - fat, ny - fat}};
    const int64_t nf1 = method == ELA_INTERP_LINEAR ? 8 : int64_t(8) * half_length * half_length * half_length;
    const int64_t nf = dipole ? 2 * nf1 : nf1;
    std::vector<int64_t> idx(nf);
    std::unordered_map<int64_t, int64_t> seen;
    seen.reserve(size_t(n_dev) * size_t(nf));
    int64_t nu = 0;
    for (int64_t d = 0; d < n_dev; ++d) {
      const float z = float(zc[d]), x = float(xc[d]), y = float(yc[d]);
      if (z - zMax > tol || z - zMin < -tol || x - xMax > tol || x - xMin < -tol || y - yMax > tol || y - yMin < -tol)
        return ELA_ERR_ARG;
      double* w = weight + d * nf;
      if (pole_weights(ax, method, half_length, z, x, y, 1.0, lin, snc, idx.data(), w) != ELA_OK) return ELA_ERR_ARG;
      if (dipole &&
          pole_weights(ax, method, half_length, float(zc[d] + z_shift), float(xc[d] + x_shift), float(yc[d] + y_shift),
                       -1.0, lin, snc, idx.data() + nf1, w + nf1) != ELA_OK)
        return ELA_ERR_ARG;
      for (int64_t k = 0; k < nf; ++k) {
        auto it = seen.find(idx[k]);
        int64_t u;
        if (it == seen.end()) {
          u = nu++;
          seen.emplace(idx[k], u);
          unique_pos[u] = idx[k];
        } else {
          u = it->second;
        }
        nbr[d * nf + k] = u;
      }
    }
    *n_unique = nu;
    return ELA_OK;
  });
}

extern "C" int ela_interp_linear(const ela_interp_axes* ax, int64_t n_dev, const double* zc, const double* xc,
                                 const double* yc, int64_t* nbr, double* weight, int64_t* unique_pos,
                                 int64_t* n_unique) {
  return guarded([&]() -> int {
    return ela_interp_weights(ax, n_dev, zc, xc, yc, ELA_INTERP_LINEAR, 1, 0, 0.0, 0.0, 0.0, nbr, weight, unique_pos,
                              n_unique);
  });
}
