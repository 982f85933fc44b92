"""Second, independent numpy restatement of the reference stepper and time loops
(test infrastructure only; tiny grids).

Written with array slicing instead of per-cell loops so that an indexing slip
in ``ela_oracle.cpp`` and in this file are unlikely to coincide. It follows
the same reference files:
elastic_iso_float/src/kernelsElaGpu_3D.cu:110-1226 (steppers, damping),
elastic_iso_float/src/kernelsElaGpu_3D.cu:11-104 (injection/recording),
elastic_iso_float/src/kernelsElaGpu_3D.cu:1232-1714 (Born imaging, coarse-time
interpolation, secondary-source injection, adjoint extraction),
elastic_iso_float/src/nonlinearPropElasticGpuFunctions_3D.cu:909-925, 1330-1345
(nonlinear time loops) and
elastic_iso_float/src/BornElasticGpuFunctions_3D.cu:1268-1407, 2147-2281 (Born
time loops). Arrays are [ny, nx, nz].
"""
from __future__ import annotations

import numpy as np

FAT = 4
C = np.array([1.196289062541883, -0.079752604188901, 0.009570312506634, -6.975446437140719e-04])
COMP_GRID = (1, 2, 3, 0, 0, 0, 4, 5, 6)


def _sl(a, axis, lo, n):
    """View of ``a`` with the interior of the other two axes and [lo, lo+n) on ``axis``."""
    ny, nx, nz = a.shape
    idx = [slice(FAT, ny - FAT), slice(FAT, nx - FAT), slice(FAT, nz - FAT)]
    idx[axis] = slice(lo, lo + n)
    return a[tuple(idx)]


def deriv(a, axis, plus, h):
    """D+ (plus=True) or D- of ``a`` along array axis (0=y, 1=x, 2=z) on the
    interior, coefficients C/h."""
    n = a.shape[axis] - 2 * FAT
    out = np.zeros(tuple(s - 2 * FAT for s in a.shape))
    for k in range(4):
        if plus:   # f[i+k+1] - f[i-k]
            out += (C[k] / h) * (_sl(a, axis, FAT + k + 1, n) - _sl(a, axis, FAT - k, n))
        else:      # f[i+k] - f[i-k-1]
            out += (C[k] / h) * (_sl(a, axis, FAT + k, n) - _sl(a, axis, FAT - k - 1, n))
    return out


def interior(a):
    return a[FAT:-FAT, FAT:-FAT, FAT:-FAT]


def step(p0, p1, med, dz, dx, sign=1.0):
    """p0 <- p0 + sign*M D(p1). y-derivatives use dx (reference behaviour)."""
    vx, vy, vz, sxx, syy, szz, sxz, sxy, syz = p1
    Bx, By, Bz, L2M, L, Mxz, Mxy, Myz = [interior(m) for m in med]
    Y, X, Z = 0, 1, 2
    exx = deriv(vx, X, True, dx)
    eyy = deriv(vy, Y, True, dx)
    ezz = deriv(vz, Z, True, dz)
    upd = [
        Bx * (deriv(sxx, X, False, dx) + deriv(sxy, Y, True, dx) + deriv(sxz, Z, True, dz)),
        By * (deriv(sxy, X, True, dx) + deriv(syy, Y, False, dx) + deriv(syz, Z, True, dz)),
        Bz * (deriv(sxz, X, True, dx) + deriv(syz, Y, True, dx) + deriv(szz, Z, False, dz)),
        L2M * exx + L * (eyy + ezz),
        L2M * eyy + L * (exx + ezz),
        L2M * ezz + L * (exx + eyy),
        Mxz * (deriv(vx, Z, False, dz) + deriv(vz, X, False, dx)),
        Mxy * (deriv(vx, Y, False, dx) + deriv(vy, X, False, dx)),
        Myz * (deriv(vz, Y, False, dx) + deriv(vy, Z, False, dz)),
    ]
    out = p0.copy()
    for k in range(9):
        interior(out[k])[...] += sign * upd[k]
    return out


def step_born_adj(p0, p1, med, dz, dx):
    """p0 <- p0 - D(M p1) (transposed stepper)."""
    vx, vy, vz, sxx, syy, szz, sxz, sxy, syz = p1
    Bx, By, Bz, L2M, L, Mxz, Mxy, Myz = med
    hv = [vx * Bx, vy * By, vz * Bz]
    hs = [sxx * L2M + (syy + szz) * L, syy * L2M + (sxx + szz) * L, szz * L2M + (sxx + syy) * L,
          sxz * Mxz, sxy * Mxy, syz * Myz]
    Y, X, Z = 0, 1, 2
    upd = [
        deriv(hs[0], X, False, dx) + deriv(hs[4], Y, True, dx) + deriv(hs[3], Z, True, dz),
        deriv(hs[4], X, True, dx) + deriv(hs[1], Y, False, dx) + deriv(hs[5], Z, True, dz),
        deriv(hs[3], X, True, dx) + deriv(hs[5], Y, True, dx) + deriv(hs[2], Z, False, dz),
        deriv(hv[0], X, True, dx),
        deriv(hv[1], Y, True, dx),
        deriv(hv[2], Z, True, dz),
        deriv(hv[0], Z, False, dz) + deriv(hv[2], X, False, dx),
        deriv(hv[0], Y, False, dx) + deriv(hv[1], X, False, dx),
        deriv(hv[2], Y, False, dx) + deriv(hv[1], Z, False, dz),
    ]
    out = p0.copy()
    for k in range(9):
        interior(out[k])[...] -= upd[k]
    return out


def damping_weights(shape, min_pad, alpha):
    """Per-cell sponge factor (1 outside the sponge) on the full grid."""
    ny, nx, nz = shape
    iy, ix, iz = np.meshgrid(np.arange(ny), np.arange(nx), np.arange(nz), indexing="ij")
    d = np.minimum.reduce([iz - FAT, ix - FAT, nz - iz - 1 - FAT, nx - ix - 1 - FAT, iy - FAT, ny - iy - 1 - FAT])
    w = np.ones(shape)
    if min_pad > 0:
        dd = np.arange(min_pad)
        tap = alpha + (1 - alpha) * np.cos(np.pi * (min_pad - dd) / min_pad)
        mask = (d >= 0) & (d < min_pad)
        w[mask] = tap[d[mask]]
    return w


def time_filter(sub, nts):
    ntw = (nts - 1) * sub + 1
    h = sub + 1
    i = np.arange(h)
    a = 1.0 - i / sub
    s = 1.0 / np.sqrt(ntw / nts)
    return np.concatenate([a * s, (1 - a) * s])


def nl_forward(g, med, src, spos, rpos):
    """Nonlinear forward on regular-grid traces (lists per component)."""
    shape = g.shape
    w = damping_weights(shape, g.minPad, g.alphaCos)
    f = time_filter(g.sub, g.nts)
    h = g.sub + 1
    p0 = np.zeros((9,) + shape)
    p1 = np.zeros((9,) + shape)
    rec = [np.zeros((len(rpos[COMP_GRID[c]]), g.nts)) for c in range(9)]
    for its in range(g.nts - 1):
        for it2 in range(1, g.sub + 1):
            p0 = step(p0, p1, med, g.dz, g.dx, 1.0)
            for c in range(9):
                pos = spos[COMP_GRID[c]]
                fl = p0[c].reshape(-1)
                np.add.at(fl, pos, src[c][:, its] * f[it2 - 1] + src[c][:, its + 1] * f[h + it2 - 1])
            p0 *= w
            p1 *= w
            for c in range(9):
                v = p0[c].reshape(-1)[rpos[COMP_GRID[c]]]
                rec[c][:, its] += v * f[it2]
                rec[c][:, its + 1] += v * f[h + it2]
            p0, p1 = p1, p0
    return rec


def _pad(inner):
    """Interior values -> full-grid array (zero FAT layer)."""
    out = np.zeros(tuple(n + 2 * FAT for n in inner.shape))
    interior(out)[...] = inner
    return out


def source_wavefield(g, med, src, spos):
    """Pass 1 of the Born operators: the background propagation with the
    coarse-time hat accumulation of vx, vy, vz
    (BornElasticGpuFunctions_3D.cu:1268-1316, kernelsElaGpu_3D.cu:1646-1672).
    Returns V [nts, 3, ny, nx, nz]."""
    shape = g.shape
    w = damping_weights(shape, g.minPad, g.alphaCos)
    f = time_filter(g.sub, g.nts)
    h = g.sub + 1
    p0 = np.zeros((9,) + shape)
    p1 = np.zeros((9,) + shape)
    V = np.zeros((g.nts, 3) + shape)
    for its in range(g.nts - 1):
        for it2 in range(1, g.sub + 1):
            p0 = step(p0, p1, med, g.dz, g.dx, 1.0)
            for c in range(9):
                fl = p0[c].reshape(-1)
                np.add.at(fl, spos[COMP_GRID[c]], src[c][:, its] * f[it2 - 1] + src[c][:, its + 1] * f[h + it2 - 1])
            p0 *= w
            p1 *= w
            V[its] += p0[:3] * f[it2]
            V[its + 1] += p0[:3] * f[h + it2]
            p0, p1 = p1, p0
    return V


def _velocity_derivatives(Vc, dz, dx):
    """Strain-rate combinations of one coarse velocity slice on the interior:
    exx, eyy, ezz, gxz = D-z vx + D-x vz, gxy = D-y vx + D-x vy,
    gyz = D-y vz + D-z vy (y coefficients over dx, as the reference)."""
    vx, vy, vz = Vc
    Y, X, Z = 0, 1, 2
    exx = deriv(vx, X, True, dx)
    eyy = deriv(vy, Y, True, dx)
    ezz = deriv(vz, Z, True, dz)
    gxz = deriv(vx, Z, False, dz) + deriv(vz, X, False, dx)
    gxy = deriv(vx, Y, False, dx) + deriv(vy, X, False, dx)
    gyz = deriv(vz, Y, False, dx) + deriv(vy, Z, False, dz)
    return exx, eyy, ezz, gxz, gxy, gyz


def _dv(V, its, nts):
    """V[its-1] - V[its+1] with V outside [0, nts) taken as zero."""
    lo = V[its - 1] if its >= 1 else 0.0
    hi = V[its + 1] if its + 1 <= nts - 1 else 0.0
    return lo - hi


def imaging_forward(g, V, its, dm, physics=False):
    """Secondary sources of coarse sample its (imagingElaFwdGpu_3D,
    kernelsElaGpu_3D.cu:1232-1438): 9 full-grid arrays. In reference mode the
    yz term uses the xz derivative pattern (:1419-1431)."""
    dinv = 1.0 / (2.0 * g.dts)
    exx, eyy, ezz, gxz, gxy, gyz = _velocity_derivatives(V[its], g.dz, g.dx)
    D = [interior(d) for d in dm]
    dv = _dv(V, its, g.nts)
    dv = [interior(dv[k]) if not np.isscalar(dv) else 0.0 for k in range(3)]
    div = exx + eyy + ezz
    ss = [D[k] * dv[k] * dinv for k in range(3)]
    ss += [D[3] * div + 2.0 * D[4] * exx, D[3] * div + 2.0 * D[4] * eyy, D[3] * div + 2.0 * D[4] * ezz,
           D[5] * gxz, D[6] * gxy, D[7] * (gyz if physics else gxz)]
    return np.stack([_pad(np.broadcast_to(a, exx.shape)) for a in ss])


def imaging_adjoint(g, V, its, ss, physics=False):
    """Images of coarse sample its (imagingElaAdjGpu_3D,
    kernelsElaGpu_3D.cu:1440-1641): 8 full-grid arrays to add."""
    dinv = 1.0 / (2.0 * g.dts)
    exx, eyy, ezz, gxz, gxy, gyz = _velocity_derivatives(V[its], g.dz, g.dx)
    S = [interior(a) for a in ss]
    dv = _dv(V, its, g.nts)
    dv = [interior(dv[k]) if not np.isscalar(dv) else 0.0 for k in range(3)]
    img = [S[k] * dv[k] * dinv for k in range(3)]
    img += [(S[3] + S[4] + S[5]) * (exx + eyy + ezz), 2.0 * (exx * S[3] + eyy * S[4] + ezz * S[5]),
            S[6] * gxz, S[7] * gxy, S[8] * (gyz if physics else gxz)]
    return np.stack([_pad(np.broadcast_to(a, exx.shape)) for a in img])


def born_forward(g, med, src, spos, dm, rpos, physics=False):
    """Born forward (BornElasticGpuFunctions_3D.cu:1226-1461): scattered
    field driven by the secondary sources of two coarse samples per fine step,
    recorded at the receivers."""
    shape = g.shape
    w = damping_weights(shape, g.minPad, g.alphaCos)
    f = time_filter(g.sub, g.nts)
    h = g.sub + 1
    V = source_wavefield(g, med, src, spos)
    p0 = np.zeros((9,) + shape)
    p1 = np.zeros((9,) + shape)
    rec = [np.zeros((len(rpos[COMP_GRID[c]]), g.nts)) for c in range(9)]
    ssL = imaging_forward(g, V, 0, dm, physics)
    for its in range(g.nts - 1):
        ssR = imaging_forward(g, V, its + 1, dm, physics)
        for it2 in range(1, g.sub + 1):
            p0 = step(p0, p1, med, g.dz, g.dx, 1.0)
            p0 += ssL * f[it2 - 1] + ssR * f[h + it2 - 1]
            p0 *= w
            p1 *= w
            for c in range(9):
                v = p0[c].reshape(-1)[rpos[COMP_GRID[c]]]
                rec[c][:, its] += v * f[it2]
                rec[c][:, its + 1] += v * f[h + it2]
            p0, p1 = p1, p0
        ssL = ssR
    return rec


def born_adjoint(g, med, src, spos, rec, rpos, physics=False):
    """Born adjoint (BornElasticGpuFunctions_3D.cu:2103-2336): the transposed
    stepper driven by the receiver traces, its coarse-time slices imaged
    against the background V. Returns the 8 images."""
    shape = g.shape
    w = damping_weights(shape, g.minPad, g.alphaCos)
    f = time_filter(g.sub, g.nts)
    h = g.sub + 1
    V = source_wavefield(g, med, src, spos)
    p0 = np.zeros((9,) + shape)
    p1 = np.zeros((9,) + shape)
    img = np.zeros((8,) + shape)
    ssL = np.zeros((9,) + shape)
    ssR = np.zeros((9,) + shape)
    for its in range(g.nts - 2, -1, -1):
        for it2 in range(g.sub - 1, -1, -1):
            p0 = step_born_adj(p0, p1, med, g.dz, g.dx)
            for c in range(9):
                fl = p0[c].reshape(-1)
                np.add.at(fl, rpos[COMP_GRID[c]], rec[c][:, its] * f[it2 + 1] + rec[c][:, its + 1] * f[h + it2 + 1])
            p0 *= w
            p1 *= w
            ssL += p0 * f[it2]
            ssR += p0 * f[h + it2]
            p0, p1 = p1, p0
        img += imaging_adjoint(g, V, its + 1, ssR, physics)
        ssR = ssL
        ssL = np.zeros((9,) + shape)
    img += imaging_adjoint(g, V, 0, ssR, physics)
    return img
