"""CPU tests of the window stacking used by the batched multi-shot path."""
import numpy as np
import pytest

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
