"""The oracle reproduces the committed golden fixture (tests/golden/small_fp64.npz):
nonlinear forward/adjoint, Born forward/adjoint and the adjoint-state
gradient on one small seeded problem."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.golden import make_golden as GD
from tests.problems import rel


@pytest.fixture(scope="module")
def golden():
    z = GD.load()
    return GD.inputs(z), GD.expected(z)


def test_fixture_inputs_are_consistent(golden):
    (g, med, spos, rpos, src, rec, dm), exp = golden
    assert med.shape == (8,) + g.shape and dm.shape == (8,) + g.shape
    assert [len(p) for p in spos] == [len(s) for s in (src[3], src[0], src[1], src[2], src[6], src[7], src[8])]
    assert all(len(np.unique(p)) == len(p) for p in spos + rpos)
    assert exp["grad8"].shape == (8,) + g.shape


def test_oracle_reproduces_golden(golden):
    (g, med, spos, rpos, src, rec, dm), exp = golden
    assert rel(O.lib_nl_forward(g, med, src, spos, rpos), exp["nl_fwd"]) < 1e-13
    assert rel(O.lib_nl_adjoint(g, med, rec, spos, rpos), exp["nl_adj"]) < 1e-13
    assert rel(O.lib_born_forward(g, med, src, spos, dm, rpos), exp["born_fwd"]) < 1e-13
    assert rel(np.asarray(O.lib_born_adjoint(g, med, src, spos, rec, rpos)), exp["born_adj"]) < 1e-13
    g8, gs = O.lib_gradient(g, med, src, spos, rec, rpos)
    assert rel(np.asarray(g8), exp["grad8"]) < 1e-13
    assert rel(gs, exp["grad_src"]) < 1e-13
