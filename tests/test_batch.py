"""CPU tests of the window stacking used by the batched multi-shot path."""
import numpy as np
import pytest

from elasticisolib_3d_amd import propagator as P
from elasticisolib_3d_amd.batch import Window, WindowStack


def _stack():
    return WindowStack((20, 30, 26), 10.0, 5, [Window(2, 3, 14, 12), Window(10, 8, 18, 15), Window(0, 0, 30, 26)])


def test_stack_geometry():
    ws = _stack()
    assert ws.nz == 20 + 5 + ws.zpads[1] + 8
    assert ws.y_off == [0, ws.ny_w[0], ws.ny_w[0] + ws.ny_w[1]]
    assert ws.ny == sum(ws.ny_w) and ws.nx == max(ws.nx_w)
    for k, w in enumerate(ws.windows):
        assert ws.ny_w[k] == w.ny + 2 * 5 + 8
        assert (ws.nz - 8) % 16 == 0 and (ws.nx_w[k] - 8) % 16 == 0   # block-multiple padding


def test_window_model_adjoint():
    ws = _stack()
    rng = np.random.default_rng(0)
    m = rng.standard_normal((2, 26, 30, 20))
    for k in range(3):
        wm = ws.window_model(k, m)
        g = rng.standard_normal(wm.shape)
        out = np.zeros_like(m)
        ws.window_adjoint(k, g, out)
        assert abs(np.sum(wm * g) - np.sum(m * out)) <= 1e-12 * abs(np.sum(m * out))
        # the window's interior is the model cut
        w = ws.windows[k]
        zm = ws.zpads[0]
        xm = ws.xpads[k][0]
        assert np.array_equal(wm[:, 4 + 5:4 + 5 + w.ny, 4 + xm:4 + xm + w.nx, 4 + zm:4 + zm + 20],
                              m[:, w.y0:w.y0 + w.ny, w.x0:w.x0 + w.nx, :])
        # every padded cell copies the model cell the edge-replication index map names
        iy, ix, iz = ws._cut_index(k)
        assert np.array_equal(wm, m[:, iy[:, None, None], ix[None, :, None], iz[None, None, :]])


def test_stacked_positions_and_stack():
    ws = _stack()
    rng = np.random.default_rng(1)
    per = [rng.standard_normal((ws.ny_w[k], ws.nx_w[k], ws.nz)) for k in range(3)]
    st = ws.stack(per)
    for k in range(3):
        pos = rng.integers(0, per[k].size, 50)
        sp = ws.stacked_positions(k, [pos])[0]
        assert np.array_equal(st.ravel()[sp], per[k].ravel()[pos])
        assert np.array_equal(ws.unstack(k, st), per[k])
    # beyond a narrower window the stack holds zeros
    k = int(np.argmin(ws.nx_w))
    assert np.all(st[ws.y_off[k]:ws.y_off[k] + ws.ny_w[k], ws.nx_w[k]:, :] == 0)


def test_window_cut_and_transpose_in_torch_match_numpy():
    """The torch window cut, stack and transpose of the cut (used to keep the
    batched gradient on the GPU) give the numpy methods' numbers."""
    import torch
    rng = np.random.default_rng(5)
    m = rng.standard_normal((3, 30, 34, 20))   # [.., ny, nx, nz]
    ws = WindowStack((20, 34, 30), 10.0, 5, [Window(2, 3, 16, 14), Window(10, 8, 20, 18)])
    mt = torch.from_numpy(m)
    for k in range(2):
        assert np.array_equal(ws.window_model_t(k, mt).numpy(), ws.window_model(k, m))
    per = [ws.window_model(k, m) for k in range(2)]
    assert np.array_equal(ws.stack_t([torch.from_numpy(a) for a in per], fill=1.0).numpy(), ws.stack(per, fill=1.0))
    for k in range(2):
        g = rng.standard_normal(per[k].shape)
        a = np.zeros_like(m)
        ws.window_adjoint(k, g, a)
        b = torch.zeros_like(mt)
        ws.window_adjoint_t(k, torch.from_numpy(g), b)
        assert np.allclose(b.numpy(), a, rtol=1e-13, atol=1e-13)


def test_medium_scaling_and_its_transpose_in_torch_match_the_library():
    """medium_torch.scale_medium / medium_gradient give the same numbers as
    the library's host ela_scale_medium / ela_medium_gradient (the same
    operations in the same order)."""
    import torch
    from elasticisolib_3d_amd import medium_torch as MT
    rng = np.random.default_rng(6)
    shape = (9, 11, 13)
    rho = 2000 * (1 + 0.2 * rng.random(shape))
    mu = 3e9 * (1 + 0.2 * rng.random(shape))
    lam = 4e9 * (1 + 0.2 * rng.random(shape))
    dtw = 5e-4
    med = P.scale_medium(rho, lam, mu, dtw)
    medt = MT.scale_medium(torch.from_numpy(rho), torch.from_numpy(lam), torch.from_numpy(mu), dtw).numpy()
    assert np.array_equal(medt, med)
    g8 = rng.standard_normal((8,) + shape)
    g3 = P.medium_gradient(g8, med, dtw)
    g3t = MT.medium_gradient(torch.from_numpy(g8), torch.from_numpy(med), dtw).numpy()
    assert np.allclose(g3t, g3, rtol=1e-14, atol=0)


def test_stack_in_a_larger_shared_grid():
    """A stack placed in a larger grid (the propagator several stacks share):
    flat positions, stacking and unstacking use the grid's extents, the
    planes past the last window and the columns past a window's width hold
    the fill value, and a grid smaller than the stack is refused."""
    natural = _stack()
    ws = WindowStack((20, 30, 26), 10.0, 5, natural.windows, grid=(natural.nx + 16, natural.ny + 9))
    assert (ws.nx, ws.ny) == (natural.nx + 16, natural.ny + 9)
    assert ws.y_off == natural.y_off and ws.ny_w == natural.ny_w and ws.nx_w == natural.nx_w
    rng = np.random.default_rng(2)
    per = [rng.standard_normal((ws.ny_w[k], ws.nx_w[k], ws.nz)) for k in range(3)]
    st = ws.stack(per, fill=1.0)
    assert st.shape == (ws.ny, ws.nx, ws.nz)
    assert np.all(st[natural.ny:] == 1.0) and np.all(st[:, natural.nx:] == 1.0)
    for k in range(3):
        pos = rng.integers(0, per[k].size, 50)
        sp = ws.stacked_positions(k, [pos])[0]
        assert np.array_equal(st.ravel()[sp], per[k].ravel()[pos])
        assert np.array_equal(ws.unstack(k, st), per[k])
    with pytest.raises(ValueError):
        WindowStack((20, 30, 26), 10.0, 5, natural.windows, grid=(natural.nx - 1, natural.ny))
