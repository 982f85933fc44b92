"""The pySolver operator surface on the GPU: the shot operators through the
C ABI against the same operators on the oracle-backed stand-in, and their
dot-product tests in FP32 and FP64."""
import numpy as np
import pytest

from elasticisolib_3d_amd import operators as OP
from elasticisolib_3d_amd.sep import Axis, Hypercube, getSepVector
from tests import operator_problem as PB
from tests.problems import rel

pytestmark = pytest.mark.gpu


def _setup(nexp=2, **extra):
    par = PB.make_par(nexp=nexp, **extra)
    ep = PB.elastic_vector()
    sgeo, rgeo = PB.geometries(nexp=nexp)
    src = OP.buildSourceGeometry_3D(par, ep, sgeo)
    rec = OP.buildReceiversGeometry_3D(par, ep, rgeo)
    t = Axis(par.getInt("nts"), 0.0, par.getFloat("dts"))
    return par, ep, src, rec, t


@pytest.mark.parametrize("cls,storage,tol", [(OP.nonlinearPropElasticShotsGpu_3D, "dataFloat", 1e-5),
                                             (OP.nonlinearPropElasticShotsGpu_3D_double, "dataDouble", 1e-12)])
def test_nonlinear_operator_dot_test(cls, storage, tol):
    par, ep, src, rec, t = _setup()
    model = getSepVector(Hypercube([t, Axis(1), Axis(9), Axis(2)]), storage=storage)
    data = getSepVector(Hypercube([t, rec[7], Axis(9), src[7]]), storage=storage)
    op = cls(model, data, ep, par, *src[:7], *rec[:7])
    (a, b, e1), (c, d, e2) = op._dot_errors(seed=3)
    print(f"{cls.__name__}: dot errors {e1:.2e} {e2:.2e}", flush=True)
    assert e1 < tol and e2 < tol
    assert op.dotTestCpp(maxError=max(tol, 1e-5), seed=3)


@pytest.mark.parametrize("cls,storage,tol", [(OP.BornElasticShotsGpu_3D, "dataFloat", 1e-5),
                                             (OP.BornElasticShotsGpu_3D_double, "dataDouble", 1e-12)])
def test_born_operator_dot_test(cls, storage, tol):
    par, ep, src, rec, t = _setup()
    sig = getSepVector(Hypercube([t, Axis(1), Axis(9)]), storage=storage)
    sig.getNdArray()[3:6] = np.sin(np.arange(par.getInt("nts")) * 0.7)
    model = getSepVector(ep.getHyper(), storage=storage)
    data = getSepVector(Hypercube([t, rec[7], Axis(9), src[7]]), storage=storage)
    op = cls(model, data, ep, par, [sig], *src[:7], *rec[:7])
    (a, b, e1), (c, d, e2) = op._dot_errors(seed=4)
    print(f"{cls.__name__}: dot errors {e1:.2e} {e2:.2e}", flush=True)
    assert e1 < tol and e2 < tol


@pytest.mark.parametrize("extra", [dict(), dict(spaceInterpMethod="sinc", hFilter1d=2, dipole=1,
                                               zDipoleShift=0.5 * PB.D, xDipoleShift=PB.D)])
def test_nonlinear_operator_matches_oracle_backend(monkeypatch, extra):
    """The GPU shot operator and the same operator on the FP64 oracle agree
    (linear devices, and sinc dipoles whose 128 support points per device
    overlap between the two poles)."""
    par, ep, src, rec, t = _setup(**extra)
    model = getSepVector(Hypercube([t, Axis(1), Axis(9), Axis(2)]), storage="dataDouble")
    model.rand(11)
    data = getSepVector(Hypercube([t, rec[7], Axis(9), src[7]]), storage="dataDouble")
    gpu = OP.nonlinearPropElasticShotsGpu_3D_double(model, data, ep, par, *src[:7], *rec[:7])
    gpu.forward(False, model, data)
    ref = data.clone()
    PB.use_oracle_backend(monkeypatch)
    cpu = OP.nonlinearPropElasticShotsGpu_3D_double(model, data, ep, par, *src[:7], *rec[:7])
    cpu.forward(False, model, ref)
    err = rel(data.getNdArray(), ref.getNdArray())
    print(f"operator GPU FP64 vs oracle {extra}: {err:.2e}", flush=True)
    assert err < 1e-11
