"""Second, independent numpy restatement of the reference stepper and time loops
(test infrastructure only; tiny grids).

Written with array slicing instead of per-cell loops so that an indexing slip
in ``ela_oracle.cpp`` and in this file are unlikely to coincide. It follows
the same reference files:
elastic_iso_float/src/kernelsElaGpu_3D.cu:110-1226 (steppers, damping),
elastic_iso_float/src/kernelsElaGpu_3D.cu:11-104 (injection/recording) and
elastic_iso_float/src/nonlinearPropElasticGpuFunctions_3D.cu:909-925, 1330-1345
(time loops). Arrays are [ny, nx, nz].
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
