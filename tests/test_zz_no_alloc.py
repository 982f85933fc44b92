"""The allocation-free nonlinear forward (ela_nl_forward_noalloc) through the
C ABI against ela_nl_forward (device memory) on the same handle: two calls in
a row, the second with fewer sources and receivers, in FP32 and FP64; and the
size limit of its scratch."""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O
from tests.problems import interior_points, random_problem, rel

pytestmark = pytest.mark.gpu

G = O.Grid(nz=29, nx=31, ny=45, nts=5, sub=8, dz=10.0, dx=10.0, dy=10.0, dts=0.004, minPad=5, alphaCos=0.95)


def _noalloc(prop, spos, src, rpos):
    """Call ela_nl_forward_noalloc on device tensors; returns the 9 receiver traces."""
    import torch
    from elasticisolib_3d_amd import _lib
    mem = _lib.ELA_MEM_DEVICE
    sp, rp = prop._points(spos, mem), prop._points(rpos, mem)
    rc = prop._counts(rpos)
    rec = prop._alloc_traces(rc, True)
    st, rt = prop._traces(src, prop._counts(spos)), prop._traces(rec, rc)
    code = prop.lib.ela_nl_forward_noalloc(prop.h, ctypes.byref(sp), ctypes.byref(st), ctypes.byref(rp),
                                           ctypes.byref(rt))
    torch.cuda.synchronize()
    prop._keep.clear()
    return code, rec


def _dev(P, dtype):
    import torch
    spos = [torch.tensor(p, device="cuda") for p in P["spos"]]
    rpos = [torch.tensor(p, device="cuda") for p in P["rpos"]]
    src = [torch.tensor(s, dtype=dtype, device="cuda") for s in P["src"]]
    return spos, src, rpos


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_noalloc_matches_nl_forward_twice(precision):
    import torch
    from elasticisolib_3d_amd.propagator import Propagator
    dtype = torch.float32 if precision == "f32" else torch.float64
    prop = Propagator(G.nz, G.nx, G.ny, G.nts, G.sub, G.dz, G.dx, G.dy, G.dts, G.minPad, G.alphaCos, precision)
    big = random_problem(G, seed=41, nsrc=4, nrec=9)
    small = random_problem(G, seed=42, nsrc=2, nrec=3)
    prop.set_medium(big["med"])
    for name, P in (("first", big), ("second, smaller", small)):
        spos, src, rpos = _dev(P, dtype)
        print(f"noalloc {precision} {name} call: {sum(len(p) for p in P['rpos'])} receiver points", flush=True)
        ref = prop.nl_forward(spos, src, rpos)
        torch.cuda.synchronize()
        ref = [r.cpu().numpy() for r in ref]
        code, out = _noalloc(prop, spos, src, rpos)
        assert code == 0, prop.lib.ela_last_error()
        out = [o.cpu().numpy() for o in out]
        e = rel(out, ref)
        same = all(np.array_equal(a, b) for a, b in zip(out, ref))
        print(f"noalloc {precision} {name}: rel {e:.2e}, bit-identical {same}", flush=True)
        assert e < (5e-6 if precision == "f32" else 1e-11)
        assert np.isfinite(np.concatenate([o.ravel() for o in out])).all()
        # and the oracle agrees with both
        orc = O.lib_nl_forward(G, P["med"] if name == "first" else big["med"], P["src"], P["spos"], P["rpos"])
        assert rel(out, orc) < (5e-6 if precision == "f32" else 1e-11)


def test_noalloc_rejects_oversized_scratch():
    """nrec_total x nts values beyond the 64 MiB scratch: ELA_ERR_ARG before any propagation."""
    import torch
    from elasticisolib_3d_amd.propagator import Propagator
    g = O.Grid(nz=29, nx=31, ny=45, nts=2000, sub=2, dz=10.0, dx=10.0, dy=10.0, dts=0.004, minPad=5, alphaCos=0.95)
    prop = Propagator(g.nz, g.nx, g.ny, g.nts, g.sub, g.dz, g.dx, g.dy, g.dts, g.minPad, g.alphaCos, "f32")
    rng = np.random.default_rng(5)
    P = random_problem(g, seed=43, nsrc=1, nrec=1)
    prop.set_medium(P["med"])
    rpos = [torch.tensor(interior_points(g, 1200, rng), device="cuda") for _ in range(7)]
    spos = [torch.tensor(p, device="cuda") for p in P["spos"]]
    src = [torch.zeros((len(P["spos"][O.COMP_GRID[c]]), g.nts), dtype=torch.float32, device="cuda") for c in range(9)]
    code, _ = _noalloc(prop, spos, src, rpos)
    msg = prop.lib.ela_last_error()
    msg = msg.decode() if isinstance(msg, bytes) else str(msg)
    print(f"oversized call: code {code}: {msg}", flush=True)
    assert code == 1
    assert "scratch" in msg
