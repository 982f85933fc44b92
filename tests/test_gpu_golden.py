"""The HIP path through the C ABI against the committed golden fixture
(tests/golden/small_fp64.npz, made by the FP64 oracle): FP64 to 1e-11 and
FP32 to 5e-6 relative L2 for the four operators and the gradient (half the
1e-5 bar for FP32 forward data)."""
import numpy as np
import pytest

from tests.golden import make_golden as GD
from tests.problems import rel

pytestmark = pytest.mark.gpu

TOL = {"f64": 1e-11, "f32": 5e-6}


@pytest.fixture(scope="module")
def golden():
    z = GD.load()
    return GD.inputs(z), GD.expected(z)


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_operators_match_golden(golden, precision):
    from elasticisolib_3d_amd.propagator import Propagator
    (g, med, spos, rpos, src, rec, dm), exp = golden
    prop = Propagator(g.nz, g.nx, g.ny, g.nts, g.sub, g.dz, g.dx, g.dy, g.dts, g.minPad, g.alphaCos, precision)
    prop.set_medium(med)
    errs = {
        "nl_fwd": rel(prop.nl_forward(spos, src, rpos), exp["nl_fwd"]),
        "nl_adj": rel(prop.nl_adjoint(spos, rec, rpos), exp["nl_adj"]),
        "born_fwd": rel(prop.born_forward(spos, src, dm, rpos), exp["born_fwd"]),
        "born_adj": rel(prop.born_adjoint(spos, src, rec, rpos), exp["born_adj"]),
    }
    prop.set_gradient_checkpoints(4, 0, "native")
    prop.gradient_forward(spos, src, rpos)
    g8, gs = prop.gradient_backward(spos, src, rpos, rec)
    errs["grad8"] = rel(g8, exp["grad8"])
    errs["grad_src"] = rel(gs, exp["grad_src"])
    print(f"golden {precision}: " + ", ".join(f"{k} {v:.2e}" for k, v in errs.items()), flush=True)
    assert all(v < TOL[precision] for v in errs.values()), errs
