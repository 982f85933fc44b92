"""A small parameter-file problem for the pySolver-surface operators, and an
oracle-backed stand-in for the GPU propagator (tests only: the stand-in lets
the CPU suite check the operators' vector layouts and plumbing)."""
from __future__ import annotations

import numpy as np

from oracle import oracle as O
from elasticisolib_3d_amd.sep import Axis, Hypercube, ParObject, Vector

NZ, NX, NY = 32, 32, 30        # (n - 2 FAT) multiple of blockSize 8 for z and x
D = 10.0
PADS = dict(zPadMinus=5, zPadPlus=5, xPadMinus=5, xPadPlus=5, yPad=5)


def make_par(nexp=2, nrec=5, nts=8, **extra):
    kw = dict(nts=nts, dts=0.004, sub=4, nz=NZ, nx=NX, ny=NY, dz=D, dx=D, dy=D, mod_par=0, fMax=20,
              blockSize=8, fat=4, alphaCos=0.95, nExp=nexp, nReceiverPerShot=nrec, iGpu=0)
    kw.update(PADS)
    kw.update(extra)
    return ParObject([], **kw)


def elastic_vector(seed=0):
    rng = np.random.default_rng(seed)
    shape = (NY, NX, NZ)
    rho = 2000.0 * (1 + 0.1 * rng.random(shape))
    vp = 2500.0 * (1 + 0.1 * rng.random(shape))
    vs = vp / np.sqrt(3.0) * (0.9 + 0.1 * rng.random(shape))
    mu = rho * vs ** 2
    lam = rho * vp ** 2 - 2 * mu
    o = -(4 + 5) * D
    hyp = Hypercube([Axis(NZ, o, D), Axis(NX, o, D), Axis(NY, o, D), Axis(3)])
    return Vector(hyp, np.stack([rho, lam, mu]).astype(np.float32))


def geometries(nexp=2, nrec=5, seed=1):
    """Source geometry [3, 1, nExp] and constant receiver geometry [3, nrec, 1]
    (rows x, y, z) inside the unpadded model."""
    rng = np.random.default_rng(seed)
    model_extent = (NZ - 8 - 10 - 1) * D
    src = np.zeros((3, 1, nexp))
    src[0, 0] = rng.uniform(2 * D, model_extent - 2 * D, nexp)
    src[1, 0] = rng.uniform(2 * D, (NY - 8 - 10 - 1) * D - 2 * D, nexp)
    src[2, 0] = rng.uniform(3 * D, model_extent - 3 * D, nexp)
    rec = np.zeros((3, nrec, 1))
    rec[0, :, 0] = rng.uniform(D, model_extent - 2 * D, nrec)
    rec[1, :, 0] = rng.uniform(D, (NY - 8 - 10 - 1) * D - 2 * D, nrec)
    rec[2, :, 0] = rng.uniform(0, 2 * D, nrec)
    return src, rec


class OracleProp:
    """CPU stand-in with the Propagator methods the shot operators call,
    computed by the FP64 oracle."""

    def __init__(self, fd):
        self.g = O.Grid(fd.nz, fd.nx, fd.ny, fd.nts, fd.sub, fd.dz, fd.dx, fd.dy, fd.dts, fd.min_pad, fd.alpha_cos)
        self.med = None
        self.physics = False

    def set_medium(self, med8):
        self.med = np.asarray(med8, np.float64)

    def set_born_mode(self, physics):
        self.physics = bool(physics)

    def nl_forward(self, spos, src, rpos):
        return O.lib_nl_forward(self.g, self.med, src, spos, rpos)

    def nl_adjoint(self, spos, rec, rpos):
        return O.lib_nl_adjoint(self.g, self.med, rec, spos, rpos)

    def born_forward(self, spos, src, dm8, rpos):
        return O.lib_born_forward(self.g, self.med, src, spos, dm8, rpos, physics=self.physics)

    def born_adjoint(self, spos, src, rec, rpos, dm8=None):
        img = O.lib_born_adjoint(self.g, self.med, src, spos, rec, rpos, physics=self.physics)
        return img if dm8 is None else dm8 + img


def use_oracle_backend(monkeypatch):
    from elasticisolib_3d_amd import operators as OP

    def make(fd, precision, device=0):
        p = OracleProp(fd)
        p.set_medium(fd.med8)
        return p

    monkeypatch.setattr(OP.FdParam, "make_propagator", make)
