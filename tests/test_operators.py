"""CPU tests of the pySolver operator surface: parameter and SEP files, the
fdParam QC, geometry builders, parameter conversion, and the shot operators'
layouts and adjointness (propagation through the oracle-backed stand-in)."""
import os

import numpy as np
import pytest

from elasticisolib_3d_amd import operators as OP
from elasticisolib_3d_amd.sep import Axis, Hypercube, ParObject, Vector, getSepVector, read_vector, write_vector
from tests import operator_problem as PB


def test_par_object(tmp_path):
    f = tmp_path / "p.par"
    f.write_text("nts=10 dts=0.004 # comment\nname=\"a b\"\niGpu=0,1\n")
    p = ParObject(["prog", f"par={f}", "nts=12", "sub=3.0"])
    assert p.getInt("nts") == 12 and p.getFloat("dts") == 0.004 and p.getString("name") == "a b"
    assert p.getInts("iGpu") == [0, 1] and p.getInt("sub") == 3 and p.getInt("absent", 7) == 7
    with pytest.raises(KeyError):
        p.getInt("absent")


def test_sep_roundtrip(tmp_path):
    hyp = Hypercube([Axis(5, 0.5, 2.0), Axis(3), Axis(2, 1.0, 0.1)])
    v = Vector(hyp, np.arange(30, dtype=np.float32).reshape(2, 3, 5))
    write_vector(str(tmp_path / "v.H"), v)
    w = read_vector(str(tmp_path / "v.H"), ndims=4)
    assert w.getNdArray().shape == (1, 2, 3, 5)
    assert np.array_equal(w.getNdArray()[0], v.getNdArray())
    assert w.getHyper().getAxis(1).o == 0.5 and w.getHyper().getAxis(3).d == pytest.approx(0.1)
    # big-endian (xdr) input
    (tmp_path / "b.bin").write_bytes(np.arange(4, dtype=">f4").tobytes())
    (tmp_path / "b.H").write_text(f'n1=4 esize=4 data_format="xdr_float" in="{tmp_path / "b.bin"}"\n')
    assert np.array_equal(read_vector(str(tmp_path / "b.H")).getNdArray(), np.arange(4, dtype=np.float32))


def test_fdparam_qc_and_staggering():
    par = PB.make_par()
    ep = PB.elastic_vector()
    fd = OP.FdParam(ep, par)
    assert fd.sub == 8 and fd.min_pad == 5 and fd.courant < 0.45
    from elasticisolib_3d_amd.propagator import scale_medium
    e = ep.getNdArray().astype(np.float64)
    assert np.allclose(fd.med8, scale_medium(e[0], e[1], e[2], fd.dtw))
    with pytest.raises(ValueError, match="Courant"):
        OP.FdParam(ep, PB.make_par(sub=1))
    with pytest.raises(ValueError, match="Dispersion"):
        OP.FdParam(ep, PB.make_par(fMax=200))
    with pytest.raises(ValueError, match="block size"):
        OP.FdParam(ep, PB.make_par(blockSize=16))
    with pytest.raises(ValueError, match="nz not consistent"):
        OP.FdParam(ep, PB.make_par(nz=PB.NZ + 1))
    with pytest.raises(ValueError, match="SUB_MAX"):
        OP.FdParam(ep, PB.make_par(sub=51))
    # freeSurface=1 takes the sponge width from zPadPlus and the x pads only
    # (fdParamElastic_3D.cpp:44-45, 88); no other change
    fs = OP.FdParam(ep, PB.make_par(freeSurface=1, zPadMinus=4, zPadPlus=6, xPadMinus=6, xPadPlus=6))
    assert fs.min_pad == 6
    assert OP.FdParam(ep, PB.make_par(zPadMinus=4, zPadPlus=6, xPadMinus=6, xPadPlus=6)).min_pad == 4
    with pytest.raises(ValueError):
        OP.FdParam(ep, PB.make_par(freeSurface=2))
    # without mod_par the C++ default (2) multiplies spacings by 1000: the
    # metre-valued hypercube then fails the consistency check
    par2 = PB.make_par()
    par2._p.pop("mod_par")
    with pytest.raises(ValueError, match="dz not consistent"):
        OP.FdParam(ep, par2)


def test_geometry_builders_and_layouts():
    par = PB.make_par()
    ep = PB.elastic_vector()
    sgeo, rgeo = PB.geometries()
    src = OP.buildSourceGeometry_3D(par, ep, sgeo)
    rec = OP.buildReceiversGeometry_3D(par, ep, rgeo)
    assert len(src) == 8 and len(src[0]) == 2 and src[7].n == 2
    assert len(rec) == 8 and len(rec[0]) == 1 and rec[7].n == 5
    # the x grid is shifted by half a cell in x only
    cx = src[0][0].getRegPosUnique()
    assert len(cx) == 8 and src[1][0].getNDeviceIrreg() == 1
    with pytest.raises(ValueError):
        OP.buildReceiversGeometry_3D(PB.make_par(nReceiverPerShot=4), ep, rgeo)
    sinc = OP.buildSourceGeometry_3D(PB.make_par(spaceInterpMethod="sinc", hFilter1d=2), ep, sgeo)
    assert sinc[0][0].weight.shape == (1, 64) and sinc[0][0].getNDeviceReg() == 64
    dip = OP.buildReceiversGeometry_3D(PB.make_par(dipole=1, zDipoleShift=0.5 * PB.D), ep, rgeo)
    assert dip[0][0].weight.shape == (5, 16)
    assert np.allclose(dip[0][0].weight.sum(axis=1), 0.0, atol=1e-6)
    with pytest.raises(ValueError):
        OP.buildSourceGeometry_3D(PB.make_par(spaceInterpMethod="cubic"), ep, sgeo)


def test_elastic_conversion():
    rng = np.random.default_rng(2)
    hyp = Hypercube([Axis(4), Axis(3), Axis(2), Axis(3)])
    m = Vector(hyp, np.stack([2500 + rng.random((2, 3, 4)), 1400 + rng.random((2, 3, 4)),
                              2000 + rng.random((2, 3, 4))]).astype(np.float64))
    d = m.cloneSpace()
    OP.ElasticConv_3D(m, 1).forward(False, m, d)
    back = m.cloneSpace()
    OP.ElasticConv_3D(m, -1).forward(False, d, back)
    assert np.allclose(back.getNdArray(), m.getNdArray())
    jab = OP.ElasticConvJab_3D(m, None, 1)
    jab.setBackground(m)
    assert jab.dotTestCpp(maxError=1e-12, seed=3)
    # linearisation: central difference of the conversion
    dm = m.clone().rand(4)
    h = 1e-3
    plus, minus = m.cloneSpace(), m.cloneSpace()
    OP.ElasticConv_3D(m, 1).forward(False, Vector(hyp, m.getNdArray() + h * dm.getNdArray()), plus)
    OP.ElasticConv_3D(m, 1).forward(False, Vector(hyp, m.getNdArray() - h * dm.getNdArray()), minus)
    lin = m.cloneSpace()
    jab.forward(False, dm, lin)
    fd = (plus.getNdArray() - minus.getNdArray()) / (2 * h)
    assert np.allclose(lin.getNdArray(), fd, rtol=1e-6, atol=1e-6 * np.abs(fd).max())


def _nl_op(monkeypatch, const_src=True, **extra):
    PB.use_oracle_backend(monkeypatch)
    par = PB.make_par(**extra)
    ep = PB.elastic_vector()
    sgeo, rgeo = PB.geometries()
    src = OP.buildSourceGeometry_3D(par, ep, sgeo)
    rec = OP.buildReceiversGeometry_3D(par, ep, rgeo)
    t = Axis(par.getInt("nts"), 0.0, par.getFloat("dts"))
    model = getSepVector(Hypercube([t, Axis(1), Axis(9), Axis(1 if const_src else 2)]), storage="dataDouble")
    data = getSepVector(Hypercube([t, rec[7], Axis(9), src[7]]), storage="dataDouble")
    op = OP.nonlinearPropElasticShotsGpu_3D(model, data, ep, par, *src[:7], *rec[:7])
    return op, model, data, ep, par, src, rec


@pytest.mark.parametrize("const_src", [True, False])
def test_nonlinear_shots_operator_dot_product(monkeypatch, const_src):
    op, model, data, *_ = _nl_op(monkeypatch, const_src)
    assert op.dotTestCpp(maxError=1e-10, seed=5)
    op.dotTest(maxError=1e-10, seed=6)


@pytest.mark.parametrize("extra", [dict(spaceInterpMethod="sinc", hFilter1d=2),
                                   dict(dipole=1, zDipoleShift=0.5 * PB.D, xDipoleShift=-0.25 * PB.D),
                                   dict(spaceInterpMethod="sinc", hFilter1d=1, dipole=1, yDipoleShift=PB.D)])
def test_nonlinear_shots_operator_dot_product_sinc_and_dipole(monkeypatch, extra):
    """Sinc and dipole devices keep the operator pair adjoint (their weights
    enter both directions through the same sparse matrix)."""
    op, model, data, *_ = _nl_op(monkeypatch, False, **extra)
    assert op.dotTestCpp(maxError=1e-10, seed=5)


def test_nonlinear_shots_operator_matches_single_shot(monkeypatch):
    """Shot 1 of the operator's data equals a direct single-shot propagation
    built from the same geometry (layout [nExp, 9, nRec, nts])."""
    from elasticisolib_3d_amd.acquisition import ShotGeometry, scale_regular_traces
    op, model, data, ep, par, src, rec = _nl_op(monkeypatch, False)
    model.rand(7)
    op.forward(False, model, data)
    sg = ShotGeometry.from_grids([src[g][1] for g in range(7)])
    rg = ShotGeometry.from_grids([rec[g][0] for g in range(7)])
    m = model.getNdArray()[1]
    reg = scale_regular_traces(op.fd.med8, sg, sg.to_regular(list(m)), PB.D ** 3)
    out = rg.to_irregular(op.prop.nl_forward(sg.positions(), reg, rg.positions()))
    assert np.allclose(np.stack(out), data.getNdArray()[1], rtol=1e-12, atol=1e-14 * np.abs(out).max())
    # add=True accumulates
    d2 = data.clone()
    op.forward(True, model, d2)
    assert np.allclose(d2.getNdArray(), 2 * data.getNdArray())


def test_born_shots_operator_dot_product(monkeypatch):
    PB.use_oracle_backend(monkeypatch)
    par = PB.make_par()
    ep = PB.elastic_vector()
    sgeo, rgeo = PB.geometries()
    src = OP.buildSourceGeometry_3D(par, ep, sgeo)
    rec = OP.buildReceiversGeometry_3D(par, ep, rgeo)
    t = Axis(par.getInt("nts"), 0.0, par.getFloat("dts"))
    sig = getSepVector(Hypercube([t, Axis(1), Axis(9)]), storage="dataDouble")
    sig.getNdArray()[5] = np.sin(np.arange(par.getInt("nts")))  # mzz
    model = getSepVector(ep.getHyper(), storage="dataDouble")
    data = getSepVector(Hypercube([t, rec[7], Axis(9), src[7]]), storage="dataDouble")
    op = OP.BornElasticShotsGpu_3D(model, data, ep, par, [sig], *src[:7], *rec[:7])
    assert op.dotTestCpp(maxError=1e-10, seed=8)


def test_shots_operator_refuses_domain_decomposition(monkeypatch):
    PB.use_oracle_backend(monkeypatch)
    with pytest.raises(NotImplementedError):
        OP._device_of(PB.make_par(domDec=1))


class _NearlyAdjoint(OP.Operator):
    """Identity forward; the adjoint also adds `eps` x data[0] to model[0], so
    the pair is adjoint up to a relative error of order eps / sqrt(n) for
    random vectors of n values."""

    def __init__(self, vec, eps):
        self.setDomainRange(vec, vec)
        self.eps = eps

    def forward(self, add, model, data):
        out = model.getNdArray().astype(data.dtype)
        data.getNdArray()[...] = data.getNdArray() + out if add else out

    def adjoint(self, add, model, data):
        out = data.getNdArray().astype(model.dtype)
        out.ravel()[0] += model.dtype.type(self.eps) * out.ravel()[0]
        model.getNdArray()[...] = model.getNdArray() + out if add else out


def _dot_error_in_float(a, b):
    f = np.float32
    return abs((f(a) - f(b)) / f(b))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_dot_test_precision_follows_the_vectors(dtype):
    """operator.h of the float build truncates the inner products to float and
    compares in float (elastic_iso_float/include/operator.h:65-66, :78-79); the
    double build keeps double. An FP64 error of ~1e-10 that float cannot
    resolve passes the FP32 test at maxError 1e-12 and fails the FP64 one; an
    error of ~1e-4 fails both. The expected outcome is derived from the FP64
    inner products independently of dotTestCpp's code path."""
    vec = Vector(Hypercube([Axis(1000), Axis(1000)]), dtype=dtype)
    small = _NearlyAdjoint(vec, 1e-6)
    (a, b, e1), (c, d, e2) = small._dot_errors(seed=11)
    assert 1e-12 < e1 < 1e-8 and 1e-12 < e2 < 1e-8   # FP64 sees the defect
    in_float = _dot_error_in_float(a, b) == 0 and _dot_error_in_float(c, d) == 0
    assert in_float  # float cannot (for this seed)
    assert small.dotTestCpp(maxError=1e-12, seed=11) == (dtype == np.float32)
    large = _NearlyAdjoint(vec, 1e3)
    assert not large.dotTestCpp(maxError=1e-5, seed=11)
    assert _NearlyAdjoint(vec, 0.0).dotTestCpp(maxError=0.0, seed=11)


def test_dot_test_zero_inner_product_is_not_an_exception():
    """A zero operator makes 0/0: the reference's float comparison of a NaN
    error with maxError is false, so its dotTest returns true; the FP64
    errors are NaN rather than an exception."""

    class Zero(OP.Operator):
        def __init__(self, vec):
            self.setDomainRange(vec, vec)

        def forward(self, add, model, data):
            if not add:
                data.zero()

        def adjoint(self, add, model, data):
            if not add:
                model.zero()

    for dtype in (np.float32, np.float64):
        op = Zero(Vector(Hypercube([Axis(8)]), dtype=dtype))
        (_, _, e1), (_, _, e2) = op._dot_errors(seed=1)
        assert np.isnan(e1) and np.isnan(e2)
        assert op.dotTestCpp(seed=1)


def test_data_component_sampling():
    hyp = Hypercube([Axis(7), Axis(5), Axis(9), Axis(2)])
    model = Vector(hyp, np.random.default_rng(3).standard_normal(hyp.shape))
    op = OP.ElasticDatComp_3D("p,vx,sxz", model)
    data = op.getRange().clone()
    op.forward(False, model, data)
    m = model.getNdArray()
    d = data.getNdArray()
    assert d.shape == (2, 3, 5, 7)
    assert np.allclose(d[:, 0], (m[:, 3] + m[:, 4] + m[:, 5]) / 3.0)
    assert np.array_equal(d[:, 1], m[:, 0]) and np.array_equal(d[:, 2], m[:, 6])
    assert op.dotTestCpp(maxError=1e-12, seed=4)
    with pytest.raises(ValueError):
        OP.ElasticDatComp_3D("vx,vx", model)
    with pytest.raises(ValueError):
        OP.ElasticDatComp_3D("foo,bar", model)


def test_command_line_drivers_round_trip(tmp_path, monkeypatch):
    """The nonlinear driver reads a par file and SEP files, runs forward and
    adjoint through the operator and writes SEP files (oracle backend)."""
    PB.use_oracle_backend(monkeypatch)
    ep = PB.elastic_vector()
    sgeo, rgeo = PB.geometries()
    write_vector(str(tmp_path / "elastic.H"), ep)
    write_vector(str(tmp_path / "src_geom.H"), Vector(Hypercube(ns=[2, 1, 3]), sgeo.astype(np.float32)))
    write_vector(str(tmp_path / "rec_geom.H"), Vector(Hypercube(ns=[1, 5, 3]), rgeo.astype(np.float32)))
    par = PB.make_par()
    lines = [f"{k}={v}" for k, v in par._p.items()]
    lines += [f"elasticParam={tmp_path / 'elastic.H'}", f"sourceGeomFile={tmp_path / 'src_geom.H'}",
              f"receiverGeomFile={tmp_path / 'rec_geom.H'}", "precision=f64"]
    (tmp_path / "run.par").write_text("\n".join(lines) + "\n")
    nts = par.getInt("nts")
    wav = np.zeros((1, 9, 1, nts), np.float32)
    wav[0, 5, 0] = np.sin(np.arange(nts))
    write_vector(str(tmp_path / "model.H"), Vector(Hypercube(ns=[nts, 1, 9, 1]), wav))
    args = [f"par={tmp_path / 'run.par'}", f"model={tmp_path / 'model.H'}", f"data={tmp_path / 'data.H'}"]
    assert OP.main_nonlinear(args) == 0
    d = read_vector(str(tmp_path / "data.H"))
    assert d.getNdArray().shape == (2, 9, 5, nts) and np.abs(d.getNdArray()).max() > 0
    assert OP.main_nonlinear(args[:1] + [f"model={tmp_path / 'back.H'}", args[2], "adj=1"]) == 0
    back = read_vector(str(tmp_path / "back.H"))
    assert back.getNdArray().shape == (1, 9, 1, nts)


def test_padding_tool(tmp_path):
    from elasticisolib_3d_amd import pad as PD
    from elasticisolib_3d_amd.models import pad_model
    rng = np.random.default_rng(0)
    m = Vector(Hypercube([Axis(20, 0.0, 10.0), Axis(25, 5.0, 10.0), Axis(7, 0.0, 10.0), Axis(3)]),
               rng.standard_normal((3, 7, 25, 20)).astype(np.float32))
    padded, p = PD.pad_vector(m, 5, 6, 3, fat=4, block=16)
    assert (p["nz"] - 8 - 0) % 16 == 0 and (p["nx"] - 8) % 16 == 0
    ref, pads = pad_model(m.getNdArray().astype(np.float64), 5, 6, 3)
    assert pads == (5, p["zPadPlus"], 6, p["xPadPlus"], 3)
    assert np.array_equal(padded.getNdArray(), ref.astype(np.float32))
    assert padded.getHyper().getAxis(2).o == 5.0 - (4 + 6) * 10.0
    fs = PD.padding(20, 25, 7, 5, 6, 3, free_surface=1)
    assert fs["zPadMinus"] == 4 and (fs["zPadMinus"] + fs["zPadPlus"] + 20) % 16 == 0
    write_vector(str(tmp_path / "m.H"), m)
    assert PD.main([f"model={tmp_path / 'm.H'}", f"data={tmp_path / 'p.H'}", "zPad=5", "xPad=6", "yPad=3"]) == 0
    assert np.array_equal(read_vector(str(tmp_path / "p.H")).getNdArray(), padded.getNdArray())


def test_fwi_outer_loop_on_a_quadratic_misfit():
    """The L-BFGS driver reduces a separable quadratic misfit (the interface
    of FwiProblem.misfit_and_gradient) and honours bounds and the mask."""
    from elasticisolib_3d_amd.fwi import invert
    rng = np.random.default_rng(0)
    shape = (3, 4, 5)
    true = [2000 + 100 * rng.random(shape), 5e9 * (1 + 0.1 * rng.random(shape)), 3e9 * (1 + 0.1 * rng.random(shape))]
    w = [1.0 / t ** 2 for t in true]

    class Quad:
        def misfit_and_gradient(self, r, l, m):
            ms = (r, l, m)
            J = 0.5 * sum(float(np.sum(wk * (mk - tk) ** 2)) for wk, mk, tk in zip(w, ms, true))
            return J, np.stack([wk * (mk - tk) for wk, mk, tk in zip(w, ms, true)])

    m0 = [t * 0.9 for t in true]
    out, hist = invert(Quad(), m0, n_iter=20)
    assert hist[-1][1] < 1e-6 * hist[0][1]
    assert all(np.allclose(o, t, rtol=1e-3) for o, t in zip(out, true))
    # bounds: rho capped below its true value
    cap = 0.95 * true[0]
    out, _ = invert(Quad(), m0, n_iter=20, bounds=((None, None, None), (cap, None, None)))
    assert np.all(out[0] <= cap * (1 + 1e-12))
    # mask: a zero mask on mu keeps it at its start
    mask = np.ones((3,) + shape)
    mask[2] = 0
    out, _ = invert(Quad(), m0, n_iter=10, mask=mask)
    assert np.allclose(out[2], m0[2])
