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
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/elastic_iso_3d.h"

namespace {

// Adjoint half-cell average along an axis of stride s and length na on a
// [ny][nx][nz] array.
void half_cell_adjoint(const double* in, double* out, int64_t n, int64_t s, int64_t na) {
  for (int64_t i = 0; i < n; ++i) {
    const int64_t a = (i / s) % na;
    double v;
    if (a == 0) v = 0.5 * in[i];
    else if (a < na - 1) v = 0.5 * (in[i] + in[i - s]);
    else v = in[i] + 0.5 * in[i - s];
    out[i] = v;
  }
}

// Forward half-cell average (the transpose of half_cell_adjoint):
//   out[i] = (in[i] + in[i+1])/2 for i < n-1,  out[n-1] = in[n-1].
void half_cell_forward(const double* in, double* out, int64_t n, int64_t s, int64_t na) {
  for (int64_t i = 0; i < n; ++i) {
    const int64_t a = (i / s) % na;
    out[i] = a < na - 1 ? 0.5 * (in[i] + in[i + s]) : in[i];
  }
}

bool dims_ok(int64_t nz, int64_t nx, int64_t ny) { return nz >= 2 && nx >= 2 && ny >= 2; }

}  // namespace

extern "C" int ela_scale_medium(int64_t nz, int64_t nx, int64_t ny, double dtw, const double* rho,
                                const double* lame, const double* mu, double* med8) {
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
  for (int64_t i = 0; i < n; ++i) {
    Bx[i] = tdt / Bx[i];
    By[i] = tdt / By[i];
    Bz[i] = tdt / Bz[i];
    L2M[i] = tdt * (lame[i] + 2.0 * mu[i]);
    L[i] = tdt * lame[i];
    Mxz[i] *= tdt;
    Mxy[i] *= tdt;
    Myz[i] *= tdt;
  }
  return ELA_OK;
}

// Born perturbations (BornElasticShotsGpu_3D.cpp:167-235): (drho, dlambda, dmu)
// -> drhox = rhoxDtw * Sx'(drho) (likewise y, z), dlame = 2 dtw dlambda,
// dmu = 2 dtw dmu, dmuxz = 2 dtw Sz'(Sx'(dmu)), dmuxy = 2 dtw Sy'(Sx'(dmu)),
// dmuyz = 2 dtw Sz'(Sy'(dmu)), with S' the half-cell average used for the
// medium (half_cell_adjoint).
extern "C" int ela_born_stagger(int64_t nz, int64_t nx, int64_t ny, double dtw, const double* med8,
                                const double* dm3, double* out8) {
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
  for (int64_t i = 0; i < n; ++i) {
    for (int k = 0; k < 3; ++k) out8[k * n + i] *= med8[k * n + i];
    out8[3 * n + i] = tdt * dlam[i];
    out8[4 * n + i] = tdt * dmu[i];
    for (int k = 5; k < 8; ++k) out8[k * n + i] *= tdt;
  }
  return ELA_OK;
}

// Adjoint of ela_born_stagger (BornElasticShotsGpu_3D.cpp:541-607): the 8
// images -> (drho, dlambda, dmu), overwritten.
extern "C" int ela_born_unstagger(int64_t nz, int64_t nx, int64_t ny, double dtw, const double* med8,
                                  const double* img8, double* out3) {
  if (!dims_ok(nz, nx, ny) || !med8 || !img8 || !out3 || !(dtw > 0)) return ELA_ERR_ARG;
  const int64_t n = nz * nx * ny, sz = 1, sx = nz, sy = nz * nx;
  const double tdt = 2.0 * dtw;
  std::vector<double> a(n), t(n), u(n);
  double* drho = out3;
  double* dlam = out3 + n;
  double* dmu = out3 + 2 * n;
  std::memset(drho, 0, n * sizeof(double));
  const int64_t st[3] = {sx, sy, sz}, na[3] = {nx, ny, nz};
  for (int k = 0; k < 3; ++k) {
    for (int64_t i = 0; i < n; ++i) a[i] = img8[k * n + i] * med8[k * n + i];
    half_cell_forward(a.data(), t.data(), n, st[k], na[k]);
    for (int64_t i = 0; i < n; ++i) drho[i] += t[i];
  }
  for (int64_t i = 0; i < n; ++i) {
    dlam[i] = tdt * img8[3 * n + i];
    dmu[i] = tdt * img8[4 * n + i];
  }
  // (outer, inner) forward averages of muxz, muxy, muyz
  const int64_t so[3][2] = {{sx, sz}, {sx, sy}, {sy, sz}};
  const int64_t no[3][2] = {{nx, nz}, {nx, ny}, {ny, nz}};
  for (int k = 0; k < 3; ++k) {
    for (int64_t i = 0; i < n; ++i) a[i] = tdt * img8[(5 + k) * n + i];
    half_cell_forward(a.data(), t.data(), n, so[k][1], no[k][1]);
    half_cell_forward(t.data(), u.data(), n, so[k][0], no[k][0]);
    for (int64_t i = 0; i < n; ++i) dmu[i] += u[i];
  }
  return ELA_OK;
}

// Gradient of a function of the 8 scaled medium arrays with respect to
// (rho, lambda, mu): the transposed Jacobian of ela_scale_medium applied to
// g8 = dJ/d(rhoxDtw ... muyzDtw). rho?Dtw = 2 dtw / S'(rho) gives
// d/drho = S(-(rho?Dtw)^2 / (2 dtw) g); lamb2MuDtw and lambDtw give
// 2 dtw (g_l2m + g_l) for lambda and 4 dtw g_l2m for mu; mu??Dtw give
// 2 dtw S_outer(S_inner(g)) for mu. out3 is overwritten.
extern "C" int ela_medium_gradient(int64_t nz, int64_t nx, int64_t ny, double dtw, const double* med8,
                                   const double* g8, double* out3) {
  if (!dims_ok(nz, nx, ny) || !med8 || !g8 || !out3 || !(dtw > 0)) return ELA_ERR_ARG;
  const int64_t n = nz * nx * ny, sz = 1, sx = nz, sy = nz * nx;
  const double tdt = 2.0 * dtw;
  std::vector<double> a(n), t(n), u(n);
  double* grho = out3;
  double* glam = out3 + n;
  double* gmu = out3 + 2 * n;
  std::memset(grho, 0, n * sizeof(double));
  const int64_t st[3] = {sx, sy, sz}, na[3] = {nx, ny, nz};
  for (int k = 0; k < 3; ++k) {
    for (int64_t i = 0; i < n; ++i) {
      const double b = med8[k * n + i];
      a[i] = -b * b / tdt * g8[k * n + i];
    }
    half_cell_forward(a.data(), t.data(), n, st[k], na[k]);
    for (int64_t i = 0; i < n; ++i) grho[i] += t[i];
  }
  for (int64_t i = 0; i < n; ++i) {
    glam[i] = tdt * (g8[3 * n + i] + g8[4 * n + i]);
    gmu[i] = 2.0 * tdt * g8[3 * n + i];
  }
  const int64_t so[3][2] = {{sx, sz}, {sx, sy}, {sy, sz}};
  const int64_t no[3][2] = {{nx, nz}, {nx, ny}, {ny, nz}};
  for (int k = 0; k < 3; ++k) {
    for (int64_t i = 0; i < n; ++i) a[i] = tdt * g8[(5 + k) * n + i];
    half_cell_forward(a.data(), t.data(), n, so[k][1], no[k][1]);
    half_cell_forward(t.data(), u.data(), n, so[k][0], no[k][0]);
    for (int64_t i = 0; i < n; ++i) gmu[i] += u[i];
  }
  return ELA_OK;
}

// ---------------------------------------------------------------------------
// Trilinear interpolation of acquisition devices onto one staggered grid: the
// linear method of spaceInterpGpu_3D (elastic_iso_float/src/spaceInterpGpu_3D.cpp:
// bounds check :177-211, weights :214-279, unique points :113-138).
//
// The fractional position is formed in single precision, as the reference
// does (float wz = (z - oz) / dz; long long zReg = wz; wz = 1 - (wz - zReg)),
// so that grid-point indices agree with the reference for coordinates that
// sit on, or within rounding of, a grid node.
// ---------------------------------------------------------------------------
#include <unordered_map>

extern "C" int ela_interp_linear(const ela_interp_axes* ax, int64_t n_dev, const double* zc, const double* xc,
                                 const double* yc, int64_t* nbr, double* weight, int64_t* unique_pos,
                                 int64_t* n_unique) {
  if (!ax || n_dev < 0 || (n_dev > 0 && (!zc || !xc || !yc || !nbr || !weight || !unique_pos || !n_unique)))
    return ELA_ERR_ARG;
  const int64_t nz = ax->nz, nx = ax->nx, ny = ax->ny, fat = ax->fat;
  // device bounds (checkOutOfBounds, spaceInterpGpu_3D.cpp:177-211)
  const float zMin = float(ax->oz + (fat + ax->z_pad_minus) * ax->dz);
  const float xMin = float(ax->ox + (fat + ax->x_pad_minus) * ax->dx);
  const float yMin = float(ax->oy + (fat + ax->y_pad) * ax->dy);
  const float zMax = zMin + float((nz - 2 * fat - ax->z_pad_minus - ax->z_pad_plus - 1) * ax->dz);
  const float xMax = xMin + float((nx - 2 * fat - ax->x_pad_minus - ax->x_pad_plus - 1) * ax->dx);
  const float yMax = yMin + float((ny - 2 * fat - 2 * ax->y_pad - 1) * ax->dy);
  const float tol = float(ax->tolerance);
  std::unordered_map<int64_t, int64_t> seen;
  seen.reserve(size_t(n_dev) * 8);
  int64_t nu = 0;
  for (int64_t d = 0; d < n_dev; ++d) {
    const float z = float(zc[d]), x = float(xc[d]), y = float(yc[d]);
    if (z - zMax > tol || z - zMin < -tol || x - xMax > tol || x - xMin < -tol || y - yMax > tol || y - yMin < -tol)
      return ELA_ERR_ARG;
    float wz = (z - float(ax->oz)) / float(ax->dz);
    float wx = (x - float(ax->ox)) / float(ax->dx);
    float wy = (y - float(ax->oy)) / float(ax->dy);
    const int64_t zr = (int64_t)wz, xr = (int64_t)wx, yr = (int64_t)wy;
    wz = 1.0f - (wz - float(zr));
    wx = 1.0f - (wx - float(xr));
    wy = 1.0f - (wy - float(yr));
    if (yr < fat + ax->y_pad || yr + 1 >= ny - fat - ax->y_pad) return ELA_ERR_ARG;
    if (xr < fat + ax->x_pad_minus || xr + 1 >= nx - fat - ax->x_pad_plus) return ELA_ERR_ARG;
    if (zr < fat + ax->z_pad_minus || zr + 1 >= nz - fat - ax->z_pad_plus) return ELA_ERR_ARG;
    const int64_t b = yr * nz * nx + xr * nz + zr;
    const int64_t off[8] = {0, 1, nz, nz + 1, nz * nx, nz * nx + 1, nz * nx + nz, nz * nx + nz + 1};
    const double az = wz, ax_ = wx, ay = wy;
    const double w[8] = {az * ax_ * ay,
                         (1.0 - az) * ax_ * ay,
                         az * (1.0 - ax_) * ay,
                         (1.0 - az) * (1.0 - ax_) * ay,
                         az * ax_ * (1.0 - ay),
                         (1.0 - az) * ax_ * (1.0 - ay),
                         az * (1.0 - ax_) * (1.0 - ay),
                         (1.0 - az) * (1.0 - ax_) * (1.0 - ay)};
    for (int k = 0; k < 8; ++k) {
      const int64_t p = b + off[k];
      auto it = seen.find(p);
      int64_t u;
      if (it == seen.end()) {
        u = nu++;
        seen.emplace(p, u);
        unique_pos[u] = p;
      } else {
        u = it->second;
      }
      nbr[d * 8 + k] = u;
      weight[d * 8 + k] = double(float(w[k]));  // the reference stores weights as float
    }
  }
  *n_unique = nu;
  return ELA_OK;
}
