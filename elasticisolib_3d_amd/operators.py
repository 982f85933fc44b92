"""pySolver operator surface of the elastic propagator.

Mirrors the reference's Python module python_float/Elastic_iso_float_3D.py
(and python_double/Elastic_iso_double_3D.py) together with the C++ shot
operators it wraps, so that driver scripts written against the reference keep
working with the MI355X library underneath:

  buildSourceGeometry_3D / buildReceiversGeometry_3D   Elastic_iso_float_3D.py:21-383
  nonlinearOpInit_3D                                   Elastic_iso_float_3D.py:389-442
  nonlinearPropElasticShotsGpu_3D                      Elastic_iso_float_3D.py:445-526,
                                                       nonlinearPropElasticShotsGpu_3D.cpp:122-502
  nonlinearFwiOpInitFloat_3D / nonlinearFwiPropElasticShotsGpu_3D
                                                       Elastic_iso_float_3D.py:530-640
  BornOpInitFloat_3D / BornElasticShotsGpu_3D          Elastic_iso_float_3D.py:646-770,
                                                       BornElasticShotsGpu_3D.cpp:143-747
  FdParam (QC, staggering, scaling)                    fdParamElastic_3D.cpp:13-195, 288-355
  ElasticConv_3D / ElasticConvJab_3D                   seis_utils/python/elasticParamConvertModule_3D.py
  ElasticDatComp_3D                                    seis_utils/python/dataCompModule_3D.py
  main_nonlinear / main_born (command-line drivers)    nonlinearPythonFloatMain_3D.py, BornPythonFloatMain_3D.py

Vectors are the numpy-backed `sep.Vector` (axis 1 fastest). Layouts are the
reference's: elastic parameters [3, ny, nx, nz] (rho, lambda, mu); nonlinear
source model [nExp or 1, 9, nSim, nts]; data [nExp, 9, nRec, nts]; Born model
= (drho, dlambda, dmu) on the elastic-parameter hypercube.

One process drives one GPU: shots run one after the other on one propagator
handle (the reference deals shots to GPUs from OpenMP threads,
nonlinearPropElasticShotsGpu_3D.cpp:178). The y-slab domain decomposition
(`domDec=1`) and the multi-GPU list are not supported and raise.
"""
from __future__ import annotations

import sys
from typing import List, Optional, Sequence

import numpy as np

from . import propagator as P
from .acquisition import COMP_GRID, PaddedAxes, ShotGeometry, SpaceInterp, scale_regular_traces
from .sep import Axis, Hypercube, ParObject, Vector, getSepVector, read_vector

SUB_MAX = 100           # varElaDeclare_3D.h:14
COURANT_MAX = 0.45      # fdParamElastic_3D.h:32
DISPERSION_MIN = 3.0    # fdParamElastic_3D.h:33


def _precision_dtype(precision: str):
    return np.float32 if precision == "f32" else np.float64


# ---------------------------------------------------------------------------
# Operator base (pyOperator.Operator subset)
# ---------------------------------------------------------------------------

class Operator:
    """forward(add, model, data) / adjoint(add, model, data); add=False
    overwrites the output, add=True accumulates (operator.h:26-27)."""

    def setDomainRange(self, domain: Vector, range_: Vector) -> None:
        self.domain = domain.cloneSpace()
        self.range = range_.cloneSpace()

    def getDomain(self) -> Vector:
        return self.domain

    def getRange(self) -> Vector:
        return self.range

    def checkDomainRange(self, model: Vector, data: Vector) -> None:
        if not self.domain.checkSame(model):
            raise ValueError("model vector does not match the operator's domain")
        if not self.range.checkSame(data):
            raise ValueError("data vector does not match the operator's range")

    def forward(self, add: bool, model: Vector, data: Vector) -> None:
        raise NotImplementedError

    def adjoint(self, add: bool, model: Vector, data: Vector) -> None:
        raise NotImplementedError

    def _dot_errors(self, seed: Optional[int] = None):
        """The dot-product test of operator.h:50-92: without add, then with add
        (outputs accumulate a second application). Inner products in FP64 (a
        zero one gives an infinite or NaN error, not an exception)."""
        d1 = self.domain.clone().rand(seed)
        r1 = self.range.clone().rand(None if seed is None else seed + 1)
        d2 = self.domain.clone()
        r2 = self.range.clone()
        self.adjoint(False, d2, r1)
        self.forward(False, d1, r2)
        dot1, dot2 = np.float64(r1.dot(r2)), np.float64(d1.dot(d2))
        self.forward(True, d1, r2)
        self.adjoint(True, d2, r1)
        dot3, dot4 = np.float64(r1.dot(r2)), np.float64(d1.dot(d2))
        with np.errstate(divide="ignore", invalid="ignore"):
            e1, e2 = abs((dot1 - dot2) / dot2), abs((dot3 - dot4) / dot4)
        return (float(dot1), float(dot2), float(e1)), (float(dot3), float(dot4), float(e2))

    def dotTest(self, verbose: bool = False, maxError: float = 1e-4, seed: Optional[int] = None) -> bool:
        """pySolver-style dot-product test: raises Warning when either relative
        error exceeds maxError."""
        (a, b, e1), (c, d, e2) = self._dot_errors(seed)
        if verbose:
            print(f"Dot-product test add=False: {a:.16e} {b:.16e} relative error {e1:.3e}")
            print(f"Dot-product test add=True:  {c:.16e} {d:.16e} relative error {e2:.3e}")
        if e1 > maxError:
            raise Warning(f"Dot products failure add=False; relative error {e1:.3e}")
        if e2 > maxError:
            raise Warning(f"Dot products failure add=True; relative error {e2:.3e}")
        return True

    def dotTestCpp(self, verb: bool = False, maxError: float = 1e-5, seed: Optional[int] = None) -> bool:
        """The C++ operator's dotTest (operator.h:50-92): returns False when
        either relative error exceeds maxError. Its precision follows the
        vectors as the reference's two builds do: on FP32 vectors the four
        inner products are truncated to float and the errors, and their
        comparison with maxError, are formed in float
        (elastic_iso_float/include/operator.h:50, :65-66, :78-79), so errors
        below float resolution read as 0; on FP64 vectors everything stays
        double (elastic_iso_double/include/operator.h). As there, only an error
        above maxError fails the test (a NaN error from a zero inner product
        does not)."""
        (a, b, e1), (c, d, e2) = self._dot_errors(seed)
        lim = maxError
        if self.range.dtype == np.float32:
            f = np.float32
            a, b, c, d = f(a), f(b), f(c), f(d)
            with np.errstate(divide="ignore", invalid="ignore"):
                e1, e2 = abs((a - b) / b), abs((c - d) / d)
            lim = f(maxError)
        if verb:
            print(f"WITHOUT ADD: dot1 {float(a):.16e} dot2 {float(b):.16e} error {float(e1):.16e}")
            print(f"WITH ADD: dot3 {float(c):.16e} dot4 {float(d):.16e} error {float(e2):.16e}")
        return not (e1 > lim or e2 > lim)


# ---------------------------------------------------------------------------
# Finite-difference parameters (fdParamElastic_3D)
# ---------------------------------------------------------------------------

class FdParam:
    """Parameter-file values, QC and the scaled medium of one elastic model
    (fdParamElastic_3D.cpp:13-195). `mod_par` defaults to 2 here as in the
    C++ class, which then scales the spacings and origins by 1000 (km input):
    the Python initialisers read mod_par with default 0 for the parameter
    conversion, so a par file without mod_par takes the km branch in C++.
    That behaviour is kept."""

    def __init__(self, elasticParam: Vector, par: ParObject):
        self.par = par
        hyp = elasticParam.getHyper()
        # freeSurface=1 only changes the sponge width rule (zPad = zPadPlus,
        # fdParamElastic_3D.cpp:44-45): the reference prints the Robertsson
        # free-surface message but its stepper applies no free-surface condition
        # and still damps towards the top face, and so does this one.
        self.free_surface = par.getInt("freeSurface", 0)
        if self.free_surface not in (0, 1):
            raise ValueError("freeSurface must be 0 or 1")
        self.nts = par.getInt("nts")
        self.dts = par.getFloat("dts", 0.0)
        self.ots = par.getFloat("ots", 0.0)
        self.sub = 2 * par.getInt("sub")
        if self.sub > SUB_MAX:
            raise ValueError(f"2*sub = {self.sub} exceeds SUB_MAX = {SUB_MAX}")
        self.ntw = (self.nts - 1) * self.sub + 1
        self.dtw = float(np.float32(self.dts) / np.float32(self.sub))
        self.nz = par.getInt("nz")
        self.z_pad_plus = par.getInt("zPadPlus")
        self.z_pad_minus = par.getInt("zPadMinus")
        self.z_pad = self.z_pad_plus if self.free_surface == 1 else min(self.z_pad_minus, self.z_pad_plus)
        self.mod_par = par.getInt("mod_par", 2)
        self.dz = par.getFloat("dz", -1.0)
        self.oz = hyp.getAxis(1).o
        self.nx = par.getInt("nx")
        self.x_pad_plus = par.getInt("xPadPlus")
        self.x_pad_minus = par.getInt("xPadMinus")
        self.x_pad = min(self.x_pad_minus, self.x_pad_plus)
        self.dx = par.getFloat("dx", -1.0)
        self.ox = hyp.getAxis(2).o
        self.ny = par.getInt("ny")
        self.y_pad = par.getInt("yPad")
        self.dy = par.getFloat("dy", -1.0)
        self.oy = hyp.getAxis(3).o
        if self.mod_par == 2:
            self.dz *= 1000.0; self.oz *= 1000.0
            self.dx *= 1000.0; self.ox *= 1000.0
            self.dy *= 1000.0; self.oy *= 1000.0
        self.f_max = par.getFloat("fMax", 1000.0)
        self.block_size = par.getInt("blockSize")
        self.fat = par.getInt("fat", 4)
        self.min_pad = min(self.z_pad, self.x_pad)
        self.alpha_cos = par.getFloat("alphaCos", 0.99)
        self.tolerance = par.getFloat("errorTolerance", 0.000001)
        ep = np.asarray(elasticParam.getNdArray())
        if ep.shape[0] != 3:
            raise ValueError("elastic parameters must have 3 components (rho, lambda, mu)")
        self._velocity_range(ep)
        self._qc(hyp)
        self.elastic = np.asarray(ep, np.float64)
        self.med8 = P.scale_medium(self.elastic[0], self.elastic[1], self.elastic[2], self.dtw)

    def _velocity_range(self, ep):
        """Min / max of Vp and non-zero Vs over [fat, n - 2 fat) on every
        axis (the reference's loop bounds, fdParamElastic_3D.cpp:95-113)."""
        f = self.fat
        sl = (slice(f, self.ny - 2 * f), slice(f, self.nx - 2 * f), slice(f, self.nz - 2 * f))
        rho = ep[0][sl].astype(np.float32)
        lam = ep[1][sl].astype(np.float32)
        mu = ep[2][sl].astype(np.float32)
        if rho.size == 0:
            self.min_vpvs, self.max_vpvs = 10000.0, -1.0
            return
        vp = np.sqrt((lam + 2 * mu) / rho)
        vs = np.sqrt(mu / rho)
        vs_nz = vs[vs != 0]
        self.min_vpvs = float(min(10000.0, vp.min(), vs_nz.min() if vs_nz.size else 10000.0))
        self.max_vpvs = float(max(-1.0, vp.max(), vs.max()))

    def _qc(self, hyp: Hypercube):
        tol = self.tolerance
        for i, (n, d, o, name) in enumerate(((self.nz, self.dz, self.oz, "z"), (self.nx, self.dx, self.ox, "x"),
                                             (self.ny, self.dy, self.oy, "y")), 1):
            ax = hyp.getAxis(i)
            if n != ax.n:
                raise ValueError(f"n{name} not consistent with parfile")
            if abs(d - ax.d) > tol:
                raise ValueError(f"d{name} not consistent with parfile")
            if abs(o - ax.o) > tol:
                raise ValueError(f"o{name} not consistent with parfile")
        min_d = min(self.dz, self.dx, self.dy)
        self.courant = self.max_vpvs * self.dtw * 2.0 / min_d
        if self.courant > COURANT_MAX:
            raise ValueError(f"Courant is too big: {self.courant} (max velocity {self.max_vpvs} m/s, "
                             f"dtw {self.dtw} s, min spacing {min_d} m)")
        self.dispersion_ratio = self.min_vpvs / (self.f_max * max(self.dz, self.dx, self.dy))
        if self.dispersion_ratio < DISPERSION_MIN:
            raise ValueError(f"Dispersion is too small: {self.dispersion_ratio} < {DISPERSION_MIN}")
        if (self.nz - 2 * self.fat) % self.block_size != 0:
            raise ValueError("nz-2xFAT not a multiple of block size")
        if (self.nx - 2 * self.fat) % self.block_size != 0:
            raise ValueError("nx-2xFAT not a multiple of block size")

    def check_time_axis(self, ax: Axis, what: str) -> None:
        """checkParfileConsistencyTime_3D (fdParamElastic_3D.cpp:329-334)."""
        if ax.n != self.nts:
            raise ValueError(f"[{what}] nts not consistent with parfile")
        if abs(self.dts - ax.d) > self.tolerance:
            raise ValueError(f"[{what}] dts not consistent with parfile")
        if abs(self.ots - ax.o) > self.tolerance:
            raise ValueError(f"[{what}] ots not consistent with parfile")

    def make_propagator(self, precision: str, device: int = 0) -> P.Propagator:
        prop = P.Propagator(self.nz, self.nx, self.ny, self.nts, self.sub, self.dz, self.dx, self.dy, self.dts,
                            self.min_pad, self.alpha_cos, precision, device)
        prop.set_medium(self.med8)
        return prop


def _device_of(par: ParObject) -> int:
    if par.getInt("domDec", 0) != 0:
        raise NotImplementedError("domDec=1 (y-slab domain decomposition over several GPUs) is not supported")
    gpus = par.getInts("iGpu", [-1])
    n_gpu = par.getInt("nGpu", -1)
    if gpus and gpus[0] >= 0:
        if len(set(gpus)) != len(gpus):
            raise ValueError("duplicate GPU ids in iGpu")
        return sorted(gpus)[0]
    if n_gpu <= 0:
        raise ValueError("Please provide a list of GPUs to be used (iGpu= or nGpu=)")
    return 0


# ---------------------------------------------------------------------------
# Acquisition geometry
# ---------------------------------------------------------------------------

class spaceInterpGpu_3D(SpaceInterp):
    """One device set on one (possibly shifted) grid: the reference's
    spaceInterpGpu_3D constructor (spaceInterpGpu_3D.cpp:13-79) with the
    linear or sinc method and optional dipoles (:214-480). The hypercube's
    origin is used as given (the builders shift it by half a cell for the
    staggered grids). The sinc support must lie inside the stepped interior,
    the dipole shifts are real numbers and the dipole's y weight uses dy; the
    reference's sinc check rejects every device, it truncates the shifts to
    integers and divides the dipole y position by dz."""

    def __init__(self, zCoord, xCoord, yCoord, hyper: Hypercube, nts: int, par: ParObject, dipole: int = 0,
                 zDipoleShift: float = 0.0, xDipoleShift: float = 0.0, yDipoleShift: float = 0.0,
                 method: str = "linear", hFilter1d: int = 1):
        if method == "linear" and hFilter1d != 1:
            raise ValueError("Half-length of interpolation filter should be set to 1 for linear interpolation")
        if method not in ("linear", "sinc"):
            raise ValueError(f"spatial interpolation method {method!r} is not supported (linear or sinc)")
        z = np.asarray(zCoord.getNdArray() if hasattr(zCoord, "getNdArray") else zCoord, np.float64).ravel()
        x = np.asarray(xCoord.getNdArray() if hasattr(xCoord, "getNdArray") else xCoord, np.float64).ravel()
        y = np.asarray(yCoord.getNdArray() if hasattr(yCoord, "getNdArray") else yCoord, np.float64).ravel()
        az, ax_, ay = hyper.getAxis(1), hyper.getAxis(2), hyper.getAxis(3)
        axes = PaddedAxes(az.n, ax_.n, ay.n, az.o, ax_.o, ay.o, az.d, ax_.d, ay.d, par.getInt("fat"),
                          par.getInt("zPadMinus"), par.getInt("zPadPlus"), par.getInt("xPadMinus"),
                          par.getInt("xPadPlus"), par.getInt("yPad"), par.getFloat("errorTolerance", 1e-6))
        super().__init__(axes, z, x, y, method, int(hFilter1d), int(dipole),
                         (float(zDipoleShift), float(xDipoleShift), float(yDipoleShift)))
        self.nt = int(nts)

    def getRegPosUnique(self) -> np.ndarray:
        return self.pos_unique

    def getNDeviceReg(self) -> int:
        return self.n_reg

    def getNDeviceIrreg(self) -> int:
        return self.n_dev


def _grid_hypers(elasticParam: Vector):
    """The 7 hypercubes of Elastic_iso_float_3D.py:82-110: centre, x, y, z,
    xz, xy, yz grids, staggered axes shifted by -d/2."""
    h = elasticParam.getHyper()
    z, x, y, p = h.getAxis(1), h.getAxis(2), h.getAxis(3), h.getAxis(4)
    zs = Axis(z.n, z.o - 0.5 * z.d, z.d)
    xs = Axis(x.n, x.o - 0.5 * x.d, x.d)
    ys = Axis(y.n, y.o - 0.5 * y.d, y.d)
    return [Hypercube([z, x, y, p]), Hypercube([z, xs, y, p]), Hypercube([z, x, ys, p]), Hypercube([zs, x, y, p]),
            Hypercube([zs, xs, y, p]), Hypercube([z, xs, ys, p]), Hypercube([zs, x, ys, p])]


def _interp_method(par: ParObject, who: str, sources: bool):
    method = par.getString("spaceInterpMethod", "linear")
    if method == "linear":
        return method, 1
    if method == "sinc":
        return method, par.getInt("hFilter1d", 1) if sources else par.getInt("hFilter1d")
    raise ValueError(f"**** ERROR [{who}]: Spatial interpolation method requested by user is not implemented ****")


def _read_geometry(par: ParObject, key: str):
    arr = read_vector(par.getString(key), ndims=3).getNdArray()
    return np.asarray(arr)  # [3 (x, y, z), nPoints, nExp]


def buildSourceGeometry_3D(parObject: ParObject, elasticParam: Vector, geometry: Optional[np.ndarray] = None):
    """7 lists (one entry per shot) of source interpolators and the shot axis
    (Elastic_iso_float_3D.py:21-191). `geometry` ([3, nSimSou, nExp], rows x,
    y, z) replaces the file named by sourceGeomFile."""
    dipole = parObject.getInt("dipole", 0)
    zs, xs, ys = (parObject.getFloat(k, 0.0) for k in ("zDipoleShift", "xDipoleShift", "yDipoleShift"))
    nts = parObject.getInt("nts")
    n_exp = parObject.getInt("nExp", -1)
    if n_exp == -1:
        raise ValueError("**** ERROR [buildSourceGeometry_3D]: User must provide the total number of shots ****")
    method, hf = _interp_method(parObject, "buildSourceGeometry_3D", True)
    geo = np.asarray(geometry) if geometry is not None else _read_geometry(parObject, "sourceGeomFile")
    if geo.shape[2] != n_exp:
        raise ValueError(f"ERROR! nExp ({n_exp}) not consistent with number of shots provided within "
                         f"sourceGeomFile ({geo.shape[2]})")
    hypers = _grid_hypers(elasticParam)
    lists: List[list] = [[] for _ in range(7)]
    for i in range(n_exp):
        z, x, y = geo[2, :, i], geo[0, :, i], geo[1, :, i]
        for g in range(7):
            lists[g].append(spaceInterpGpu_3D(z, x, y, hypers[g], nts, parObject, dipole, zs, xs, ys, method, hf))
    source_axis = Axis(n_exp, 0.0, 1.0)
    return (*lists, source_axis)


def buildReceiversGeometry_3D(parObject: ParObject, elasticParam: Vector, geometry: Optional[np.ndarray] = None):
    """7 lists of receiver interpolators (one entry, or one per shot) and the
    receiver axis (Elastic_iso_float_3D.py:194-383). `geometry`
    ([3, nReceiverPerShot, 1 or nExp]) replaces receiverGeomFile."""
    nts = parObject.getInt("nts")
    dipole = parObject.getInt("dipole", 0)
    # real shifts (the reference reads the receivers' with getInt, truncating them)
    zs, xs, ys = (parObject.getFloat(k, 0.0) for k in ("zDipoleShift", "xDipoleShift", "yDipoleShift"))
    n_exp = parObject.getInt("nExp", -1)
    if n_exp == -1:
        raise ValueError("**** ERROR [buildReceiversGeometry_3D]: User must provide the total number of shots ****")
    n_rec = parObject.getInt("nReceiverPerShot", -1)
    if n_rec == -1:
        raise ValueError("**** ERROR [buildReceiversGeometry_3D]: User must provide the total number of receivers "
                         "per shot (this number must be the same for all shots [nReceiverPerShot]) ****")
    method, hf = _interp_method(parObject, "buildReceiversGeometry_3D", False)
    geo = np.asarray(geometry) if geometry is not None else _read_geometry(parObject, "receiverGeomFile")
    if geo.shape[2] != 1 and geo.shape[2] != n_exp:
        raise ValueError(f"Number of experiments from parfile (#shot={n_exp}) not consistent with receivers' "
                         f"geometry file (#shots={geo.shape[2]})")
    if geo.shape[1] != n_rec:
        raise ValueError(f"Number of receivers from parfile (#receivers={n_rec}) not consistent with receivers' "
                         f"geometry file (#receivers={geo.shape[1]})")
    hypers = _grid_hypers(elasticParam)
    lists: List[list] = [[] for _ in range(7)]
    for i in range(geo.shape[2]):
        z, x, y = geo[2, :, i], geo[0, :, i], geo[1, :, i]
        for g in range(7):
            lists[g].append(spaceInterpGpu_3D(z, x, y, hypers[g], nts, parObject, dipole, zs, xs, ys, method, hf))
    receiver_axis = Axis(n_rec, 0.0, 1.0)
    return (*lists, receiver_axis)


def _shot_geometries(lists7: Sequence[list]) -> List[ShotGeometry]:
    n = len(lists7[0])
    if any(len(l) != n for l in lists7):
        raise ValueError("the 7 grid lists must have the same number of entries")
    return [ShotGeometry.from_grids([lists7[g][i] for g in range(7)]) for i in range(n)]


# ---------------------------------------------------------------------------
# Parameter conversion (elasticParamConvertModule_3D.py)
# ---------------------------------------------------------------------------

class ElasticConv_3D(Operator):
    """conv_type 1: (Vp, Vs, rho) m/s, kg/m3 -> (rho, lambda, mu); 2: km/s,
    g/cm3 -> kg/m3, Pa; -1: (rho, lambda, mu) -> (Vp, Vs, rho)."""

    def __init__(self, domain: Vector, conv_type: int):
        self.setDomainRange(domain, domain)
        self.conv_type = conv_type

    def forward(self, add, model, data):
        self.checkDomainRange(model, data)
        m, d = model.getNdArray(), data.getNdArray()
        if not add:
            data.zero()
        if self.conv_type == 1:
            d[0] += m[2]
            d[1] += m[2] * (m[0] * m[0] - 2.0 * m[1] * m[1])
            d[2] += m[2] * m[1] * m[1]
        elif self.conv_type == 2:
            d[0] += m[2] * 1e3
            d[1] += m[2] * (m[0] * m[0] - 2.0 * m[1] * m[1]) * 1e9
            d[2] += m[2] * m[1] * m[1] * 1e9
        elif self.conv_type == -1:
            d[0] += np.sqrt((m[1] + 2 * m[2]) / m[0])
            d[1] += np.sqrt(m[2] / m[0])
            d[2] += m[0]
        else:
            raise ValueError(f"ERROR! Unsupported conv_type (current value: {self.conv_type}). Please redefine it!")


class ElasticConvJab_3D(Operator):
    """Jacobian of ElasticConv_3D types 1 and 2 around a background. As in
    the reference, the constructor takes the background from `domain` (its
    `background` argument is not used) and setBackground(b) copies b,
    multiplied by 1e3 for conv_type 2."""

    def __init__(self, domain: Vector, background: Vector, conv_type: int):
        self.setDomainRange(domain, domain)
        self.background = domain.clone()
        self.conv_type = conv_type

    def forward(self, add, model, data):
        self.checkDomainRange(model, data)
        m, b, d = model.getNdArray(), self.background.getNdArray(), data.getNdArray()
        if not add:
            data.zero()
        if self.conv_type not in (1, 2):
            raise ValueError(f"ERROR! Unsupported conv_type (current value: {self.conv_type}). Please redefine it!")
        d[0] += m[2]
        d[1] += ((b[0] * b[0] - 2.0 * b[1] * b[1]) * m[2] - 4.0 * b[1] * b[2] * m[1] + 2.0 * b[0] * b[2] * m[0])
        d[2] += b[1] * b[1] * m[2] + 2.0 * b[1] * b[2] * m[1]

    def adjoint(self, add, model, data):
        self.checkDomainRange(model, data)
        m, b, d = model.getNdArray(), self.background.getNdArray(), data.getNdArray()
        if not add:
            model.zero()
        if self.conv_type not in (1, 2):
            raise ValueError(f"ERROR! Unsupported conv_type (current value: {self.conv_type}). Please redefine it!")
        m[0] += 2.0 * b[0] * b[2] * d[1]
        m[1] += 2.0 * b[1] * b[2] * d[2] - 4.0 * b[1] * b[2] * d[1]
        m[2] += d[0] + (b[0] * b[0] - 2.0 * b[1] * b[1]) * d[1] + b[1] * b[1] * d[2]

    def setBackground(self, background: Vector):
        self.background.copy(background)
        if self.conv_type == 2:
            self.background.getNdArray()[:] *= 1e3


class ElasticDatComp_3D(Operator):
    """Sampling of data components (seis_utils/python/dataCompModule_3D.py:11-141):
    elastic data [nExp, 9, nRec, nts] -> the listed components in the listed
    order (vx vy vz sxx syy szz sxz sxy syz and p = (sxx + syy + szz) / 3).
    As in the reference, a component given twice raises, and the list is
    rejected only when it names no known component."""

    NAMES = ("vx", "vy", "vz", "sxx", "syy", "szz", "sxz", "sxy", "syz")
    P_FACTOR = 1.0 / 3.0

    def __init__(self, components: str, domain: Vector):
        comps = components.split(",")
        if len(set(comps)) != len(comps):
            raise ValueError("ERROR! A component was provided multiple times! Check your input arguments")
        h = domain.getHyper()
        rng = Vector(Hypercube([h.getAxis(1), h.getAxis(2), Axis(len(comps), label=f"Component {components}"),
                                h.getAxis(4)]), dtype=domain.dtype)
        self.setDomainRange(domain, rng)
        self.comp_list = [c.lower() for c in comps]
        if not any(c in self.comp_list for c in self.NAMES + ("p",)):
            raise ValueError(f"ERROR! Provided unknown data components: {components}")

    def forward(self, add, model, data):
        self.checkDomainRange(model, data)
        if not add:
            data.zero()
        m, d = model.getNdArray(), data.getNdArray()
        for k, name in enumerate(self.NAMES):
            if name in self.comp_list:
                d[:, self.comp_list.index(name)] += m[:, k]
        if "p" in self.comp_list:
            d[:, self.comp_list.index("p")] += self.P_FACTOR * (m[:, 3] + m[:, 4] + m[:, 5])

    def adjoint(self, add, model, data):
        self.checkDomainRange(model, data)
        if not add:
            model.zero()
        m, d = model.getNdArray(), data.getNdArray()
        for k, name in enumerate(self.NAMES):
            if name in self.comp_list:
                m[:, k] += d[:, self.comp_list.index(name)]
        if "p" in self.comp_list:
            i = self.comp_list.index("p")
            for k in (3, 4, 5):
                m[:, k] += self.P_FACTOR * d[:, i]


# ---------------------------------------------------------------------------
# Shot operators
# ---------------------------------------------------------------------------

class _ShotsBase(Operator):
    precision = "f32"

    def _setup(self, elasticParam, paramP, src_lists, rec_lists):
        self.par = paramP
        self.n_exp = paramP.getInt("nExp")
        self.device = _device_of(paramP)
        self.src_geoms = _shot_geometries(src_lists)
        self.rec_geoms = _shot_geometries(rec_lists)
        if len(self.src_geoms) != self.n_exp:
            raise ValueError("number of source geometries differs from nExp")
        self.fd = FdParam(elasticParam, paramP)
        self.prop = self.fd.make_propagator(self.precision, self.device)
        self.dxyz = self.fd.dx * self.fd.dy * self.fd.dz

    def _rec(self, i: int) -> ShotGeometry:
        return self.rec_geoms[0] if len(self.rec_geoms) == 1 else self.rec_geoms[i]

    def setBackground(self, elasticParam: Vector) -> None:
        """New elastic parameters (rho, lambda, mu) for the next calls
        (nonlinearPropElasticShotsGpu_3D.h:53 rebuilds the fdParam)."""
        self.fd = FdParam(elasticParam, self.par)
        self.prop.set_medium(self.fd.med8)

    def _scaled_regular(self, geom: ShotGeometry, irreg9) -> List[np.ndarray]:
        return scale_regular_traces(self.fd.med8, geom, geom.to_regular(list(irreg9)), self.dxyz)


class nonlinearPropElasticShotsGpu_3D(_ShotsBase):
    """Nonlinear shots operator: model = source signals [nExp or 1, 9, nSim,
    nts] (fx fy fz mxx myy mzz mxz mxy myz), data = [nExp, 9, nRec, nts]."""

    def __init__(self, domain, range, elasticParam, paramP, sourcesVectorCenterGrid, sourcesVectorXGrid,
                 sourcesVectorYGrid, sourcesVectorZGrid, sourcesVectorXZGrid, sourcesVectorXYGrid, sourcesVectorYZGrid,
                 recVectorCenterGrid, recVectorXGrid, recVectorYGrid, recVectorZGrid, recVectorXZGrid,
                 recVectorXYGrid, recVectorYZGrid):
        self.setDomainRange(domain, range)
        self._setup(elasticParam, paramP,
                    (sourcesVectorCenterGrid, sourcesVectorXGrid, sourcesVectorYGrid, sourcesVectorZGrid,
                     sourcesVectorXZGrid, sourcesVectorXYGrid, sourcesVectorYZGrid),
                    (recVectorCenterGrid, recVectorXGrid, recVectorYGrid, recVectorZGrid, recVectorXZGrid,
                     recVectorXYGrid, recVectorYZGrid))

    def _check(self, model: Vector):
        if model.getHyper().getAxis(3).n != 9:
            raise ValueError("Number of components in source model different than 9 "
                             "(fx,fy,fz,mxx,myy,mzz,mxz,mxy,myz)")
        self.fd.check_time_axis(model.getHyper().getAxis(1), "Model File")

    def forward(self, add, model, data):
        """nonlinearPropElasticShotsGpu_3D.cpp:122-199 and
        nonlinearPropElasticGpu_3D.cpp:66-335."""
        self._check(model)
        m = model.getNdArray()
        d = data.getNdArray()
        if not add:
            data.zero()
        const_src = m.shape[0] == 1
        for i in range(self.n_exp):
            sg, rg = self.src_geoms[i], self._rec(i)
            src_reg = self._scaled_regular(sg, m[0 if const_src else i])
            rec_reg = self.prop.nl_forward(sg.positions(), src_reg, rg.positions())
            d[i] += np.stack(rg.to_irregular(rec_reg)).astype(d.dtype)

    def adjoint(self, add, model, data):
        """nonlinearPropElasticShotsGpu_3D.cpp:298-400 and
        nonlinearPropElasticGpu_3D.cpp:337-615: receiver traces are mapped to
        the grids and medium-scaled, propagated backwards, and extracted at
        the sources; with one source signal for all shots the shots stack."""
        self._check(model)
        m = model.getNdArray()
        d = data.getNdArray()
        if not add:
            model.zero()
        const_src = m.shape[0] == 1
        for i in range(self.n_exp):
            sg, rg = self.src_geoms[i], self._rec(i)
            rec_reg = self._scaled_regular(rg, d[i])
            src_reg = self.prop.nl_adjoint(sg.positions(), rec_reg, rg.positions())
            m[0 if const_src else i] += np.stack(sg.to_irregular(src_reg)).astype(m.dtype)


class nonlinearFwiPropElasticShotsGpu_3D(nonlinearPropElasticShotsGpu_3D):
    """Nonlinear operator with the elastic parameters as its model: forward
    sets them as background and models the fixed sources
    (Elastic_iso_float_3D.py:592-640)."""

    def __init__(self, domain, range, sources, paramP, *lists14):
        self.sources = sources.clone()
        super().__init__(domain, range, domain, paramP, *lists14)
        self.setDomainRange(domain, range)

    def __str__(self):
        return " Ela NLOper 3D "

    def forward(self, add, model, data):
        self.setBackground(model)
        super().forward(add, self.sources, data)

    def adjoint(self, add, model, data):
        raise NotImplementedError("the FWI operator has no adjoint: use its gradient (Born adjoint or the "
                                  "adjoint-state gradient)")


class BornElasticShotsGpu_3D(_ShotsBase):
    """Born shots operator: model = (drho, dlambda, dmu) on the elastic
    parameter hypercube, data = [nExp, 9, nRec, nts]
    (BornElasticShotsGpu_3D.cpp:143-747)."""

    def __init__(self, domain, range, elasticParam, paramP, sourcesSignalsVector, sourcesVectorCenterGrid,
                 sourcesVectorXGrid, sourcesVectorYGrid, sourcesVectorZGrid, sourcesVectorXZGrid, sourcesVectorXYGrid,
                 sourcesVectorYZGrid, recVectorCenterGrid, recVectorXGrid, recVectorYGrid, recVectorZGrid,
                 recVectorXZGrid, recVectorXYGrid, recVectorYZGrid):
        self.setDomainRange(domain, range)
        self.signals = [np.asarray(s.getNdArray() if hasattr(s, "getNdArray") else s) for s in sourcesSignalsVector]
        self._setup(elasticParam, paramP,
                    (sourcesVectorCenterGrid, sourcesVectorXGrid, sourcesVectorYGrid, sourcesVectorZGrid,
                     sourcesVectorXZGrid, sourcesVectorXYGrid, sourcesVectorYZGrid),
                    (recVectorCenterGrid, recVectorXGrid, recVectorYGrid, recVectorZGrid, recVectorXZGrid,
                     recVectorXYGrid, recVectorYZGrid))
        # par key bornPhysics=1 (not in the reference, default 0): the yz shear
        # term with the yz derivative pattern instead of the reference's xz one
        # (kernelsElaGpu_3D.cu:1419-1431), the true linearisation in mu
        self.prop.set_born_mode(paramP.getInt("bornPhysics", 0) != 0)

    def _signal(self, i: int) -> np.ndarray:
        s = self.signals[0] if len(self.signals) == 1 else self.signals[i]
        return s.reshape(9, -1, self.fd.nts)

    def forward(self, add, model, data):
        """Perturbations staggered and scaled once (BornElasticShotsGpu_3D.cpp:167-235),
        then one Born propagation per shot."""
        d = data.getNdArray()
        if not add:
            data.zero()
        dm8 = P.born_stagger(np.asarray(model.getNdArray(), np.float64), self.fd.med8, self.fd.dtw)
        for i in range(self.n_exp):
            sg, rg = self.src_geoms[i], self._rec(i)
            src_reg = self._scaled_regular(sg, self._signal(i))
            rec_reg = self.prop.born_forward(sg.positions(), src_reg, dm8, rg.positions())
            d[i] += np.stack(rg.to_irregular(rec_reg)).astype(d.dtype)

    def adjoint(self, add, model, data):
        """Images of all shots accumulated on the staggered grids, then
        un-staggered once (BornElasticShotsGpu_3D.cpp:437-607)."""
        d = data.getNdArray()
        m = model.getNdArray()
        if not add:
            model.zero()
        img8 = np.zeros((8,) + tuple(self.fd.med8.shape[1:]), _precision_dtype(self.precision))
        for i in range(self.n_exp):
            sg, rg = self.src_geoms[i], self._rec(i)
            src_reg = self._scaled_regular(sg, self._signal(i))
            rec_reg = rg.to_regular(list(d[i]))
            img8 = self.prop.born_adjoint(sg.positions(), src_reg, rec_reg, rg.positions(), img8)
        m += P.born_unstagger(np.asarray(img8, np.float64), self.fd.med8, self.fd.dtw).astype(m.dtype)


class nonlinearPropElasticShotsGpu_3D_double(nonlinearPropElasticShotsGpu_3D):
    """FP64 twin (python_double/Elastic_iso_double_3D.py:452-535)."""
    precision = "f64"


class BornElasticShotsGpu_3D_double(BornElasticShotsGpu_3D):
    """FP64 twin (python_double/Elastic_iso_double_3D.py:616-690)."""
    precision = "f64"


# ---------------------------------------------------------------------------
# Initialisers
# ---------------------------------------------------------------------------

def _elastic_param(par: ParObject, storage: str) -> Vector:
    name = par.getString("elasticParam", "noElasticParamFile")
    if name == "noElasticParamFile":
        raise IOError("**** ERROR: User did not provide elastic parameter file ****")
    v = read_vector(name, ndims=4, dtype=np.float32 if storage == "dataFloat" else np.float64)
    mod_par = par.getInt("mod_par", 0)
    if mod_par != 0:
        conv = ElasticConv_3D(v, mod_par)
        tmp = v.clone()
        conv.forward(False, tmp, v)
    return v


def _storage(precision: str) -> str:
    return "dataFloat" if precision == "f32" else "dataDouble"


def nonlinearOpInit_3D(args, precision: str = "f32"):
    """Elastic_iso_float_3D.py:389-442: model, data, elastic parameters, par
    object and the 14 geometry lists."""
    par = args if isinstance(args, ParObject) else ParObject(args)
    storage = _storage(precision)
    time_axis = Axis(par.getInt("nts"), par.getFloat("ots", 0.0), par.getFloat("dts"))
    src_geo = read_vector(par.getString("sourceGeomFile"), ndims=3)
    sim_axis = src_geo.getHyper().getAxis(2)
    model = getSepVector(Hypercube([time_axis, sim_axis, Axis(9), Axis(1)]), storage=storage)
    elastic = _elastic_param(par, storage)
    src = buildSourceGeometry_3D(par, elastic, src_geo.getNdArray())
    rec = buildReceiversGeometry_3D(par, elastic)
    data = getSepVector(Hypercube([time_axis, rec[7], Axis(9), src[7]]), storage=storage)
    return (model, data, elastic, par, *src[:7], *rec[:7])


def nonlinearFwiOpInitFloat_3D(args):
    """Elastic_iso_float_3D.py:530-589: elastic parameters (as read and as
    converted), data, sources, par object and the 14 geometry lists."""
    par = args if isinstance(args, ParObject) else ParObject(args)
    time_axis = Axis(par.getInt("nts"), par.getFloat("ots", 0.0), par.getFloat("dts"))
    src_geo = read_vector(par.getString("sourceGeomFile"), ndims=3)
    sim_axis = src_geo.getHyper().getAxis(2)
    sources = getSepVector(Hypercube([time_axis, sim_axis, Axis(9), Axis(1)]))
    sources.getNdArray().flat[:] = read_vector(par.getString("sources", "noSourcesFile"), ndims=4).getNdArray().ravel()
    name = par.getString("elasticParam", "noElasticParamFile")
    if name == "noElasticParamFile":
        raise IOError("**** ERROR: User did not provide elastic parameter file ****")
    elastic = read_vector(name, ndims=4)
    elastic_conv = elastic
    mod_par = par.getInt("mod_par", 0)
    if mod_par != 0:
        conv = ElasticConv_3D(elastic, mod_par)
        elastic_conv = elastic.clone()
        conv.forward(False, elastic.clone(), elastic_conv)
    src = buildSourceGeometry_3D(par, elastic, src_geo.getNdArray())
    rec = buildReceiversGeometry_3D(par, elastic)
    data = getSepVector(Hypercube([time_axis, rec[7], Axis(9), src[7]]))
    return (elastic, elastic_conv, data, sources, par, *src[:7], *rec[:7])


def BornOpInitFloat_3D(args, precision: str = "f32"):
    """Elastic_iso_float_3D.py:646-713: model, data, elastic parameters, par
    object, source signals and the 14 geometry lists."""
    par = args if isinstance(args, ParObject) else ParObject(args)
    storage = _storage(precision)
    elastic = _elastic_param(par, storage)
    time_axis = Axis(par.getInt("nts"), par.getFloat("ots", 0.0), par.getFloat("dts"))
    src_geo = read_vector(par.getString("sourceGeomFile"), ndims=3)
    sim_axis = src_geo.getHyper().getAxis(2)
    sources_file = par.getString("sources", "noSourcesFile")
    if sources_file == "noSourcesFile":
        raise IOError("**** ERROR: User did not provide seismic sources file ****")
    signals = getSepVector(Hypercube([time_axis, sim_axis, Axis(9)]), storage=storage)
    signals.getNdArray().flat[:] = read_vector(sources_file, ndims=4).getNdArray().ravel()
    src = buildSourceGeometry_3D(par, elastic, src_geo.getNdArray())
    rec = buildReceiversGeometry_3D(par, elastic)
    model = getSepVector(elastic.getHyper(), storage=storage)
    data = getSepVector(Hypercube([time_axis, rec[7], Axis(9), src[7]]), storage=storage)
    return (model, data, elastic, par, [signals], *src[:7], *rec[:7])


def BornOpInitDouble_3D(args):
    return BornOpInitFloat_3D(args, precision="f64")


# ---------------------------------------------------------------------------
# Command-line drivers (python -m elasticisolib_3d_amd.operators nonlinear|born
# par=... model=... data=... [adj=1] [dpTest=1] [precision=f64])
# ---------------------------------------------------------------------------

def main_nonlinear(argv) -> int:
    """nonlinearPythonFloatMain_3D.py: forward (adj=0) reads model= and writes
    data=, adjoint (adj=1) reads data= and writes model=; dpTest=1 runs the
    dot-product test."""
    from .sep import write_vector
    par = ParObject(argv)
    prec = par.getString("precision", "f32")
    model, data, elastic, par, *geom = nonlinearOpInit_3D(par, prec)
    cls = nonlinearPropElasticShotsGpu_3D if prec == "f32" else nonlinearPropElasticShotsGpu_3D_double
    op = cls(model, data, elastic, par, *geom)
    if par.getInt("dpTest", 0) == 1:
        op.dotTest(True)
        return 0
    model_file = par.getString("model", "noModelFile")
    data_file = par.getString("data", "noDataFile")
    if model_file == "noModelFile":
        raise IOError("**** ERROR: User did not provide model file [model] ****")
    if data_file == "noDataFile":
        raise IOError("**** ERROR: User did not provide data file name [data] ****")
    if par.getInt("adj", 0) == 0:
        model.getNdArray().flat[:] = read_vector(model_file, ndims=4).getNdArray().ravel()
        op.forward(False, model, data)
        write_vector(data_file, data)
    else:
        data.getNdArray().flat[:] = read_vector(data_file, ndims=4).getNdArray().ravel()
        op.adjoint(False, model, data)
        write_vector(model_file, model)
    return 0


def main_born(argv) -> int:
    """BornPythonFloatMain_3D.py: forward reads model= (drho, dlambda, dmu)
    and writes data=; adjoint reads data= and writes model=."""
    from .sep import write_vector
    par = ParObject(argv)
    prec = par.getString("precision", "f32")
    model, data, elastic, par, signals, *geom = BornOpInitFloat_3D(par, prec)
    cls = BornElasticShotsGpu_3D if prec == "f32" else BornElasticShotsGpu_3D_double
    op = cls(model, data, elastic, par, signals, *geom)
    if par.getInt("dpTest", 0) == 1:
        op.dotTest(True)
        return 0
    model_file = par.getString("model", "noModelFile")
    data_file = par.getString("data", "noDataFile")
    if model_file == "noModelFile":
        raise IOError("**** ERROR: User did not provide model file ****")
    if data_file == "noDataFile":
        raise IOError("**** ERROR: User did not provide data file name ****")
    if par.getInt("adj", 0) == 0:
        model.getNdArray().flat[:] = read_vector(model_file).getNdArray().ravel()
        op.forward(False, model, data)
        write_vector(data_file, data)
    else:
        data.getNdArray().flat[:] = read_vector(data_file, ndims=4).getNdArray().ravel()
        op.adjoint(False, model, data)
        write_vector(model_file, model)
    return 0


if __name__ == "__main__":
    if len(sys.argv) < 2 or sys.argv[1] not in ("nonlinear", "born"):
        print("usage: python -m elasticisolib_3d_amd.operators nonlinear|born par=... model=... data=... [adj=1]",
              file=sys.stderr)
        sys.exit(2)
    sys.exit(main_nonlinear(sys.argv[2:]) if sys.argv[1] == "nonlinear" else main_born(sys.argv[2:]))
