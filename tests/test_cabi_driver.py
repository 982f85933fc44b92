"""A C++ program calling the C ABI with host buffers, as the reference's
single-shot operators call their GPU layer (tests/cabi_driver.cpp): it
compiles and links against the library with g++ (no GPU needed), and on the
GPU its four operator outputs match the golden fixture."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from tests.golden import make_golden as GD
from tests.problems import rel

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "elasticisolib_3d_amd", "lib")
COMP_GRID = (1, 2, 3, 0, 0, 0, 4, 5, 6)


def _compile(out_dir) -> str:
    exe = os.path.join(str(out_dir), "cabi_driver")
    rocm_lib = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cabi_driver.cpp"), "-o", exe, "-L", LIBDIR, "-lelastic_iso_3d",
           f"-Wl,-rpath,{LIBDIR}", f"-Wl,-rpath-link,{rocm_lib}"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_cpp_caller_compiles_and_links(tmp_path):
    if not os.path.exists(os.path.join(LIBDIR, "libelastic_iso_3d.so")):
        pytest.skip("library not built")
    assert os.path.exists(_compile(tmp_path))


def _write_problem(d, g, med, spos, rpos, src, rec, dm):
    with open(os.path.join(d, "grid.txt"), "w") as f:
        f.write(f"{g.nz} {g.nx} {g.ny} {g.nts} {g.sub} {g.dz!r} {g.dx!r} {g.dy!r} {g.dts!r} {g.minPad} "
                f"{g.alphaCos!r}\n")
    np.ascontiguousarray(med, np.float64).tofile(os.path.join(d, "med.bin"))
    np.ascontiguousarray(dm, np.float64).tofile(os.path.join(d, "dm.bin"))
    for name, pos in (("spos", spos), ("rpos", rpos)):
        np.concatenate([np.array([len(p) for p in pos], np.int64)] + [np.asarray(p, np.int64) for p in pos]).tofile(
            os.path.join(d, name + ".bin"))
    for name, tr in (("src", src), ("rec", rec)):
        np.concatenate([np.ascontiguousarray(t, np.float64).ravel() for t in tr]).tofile(os.path.join(d, name + ".bin"))


def _split(flat, pos, nts):
    out, o = [], 0
    for c in range(9):
        n = len(pos[COMP_GRID[c]]) * nts
        out.append(flat[o:o + n].reshape(-1, nts))
        o += n
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol", [("f64", 1e-11), ("f32", 5e-6)])
def test_cpp_caller_matches_golden(tmp_path, precision, tol):
    exe = _compile(tmp_path)
    z = GD.load()
    g, med, spos, rpos, src, rec, dm = GD.inputs(z)
    exp = GD.expected(z)
    _write_problem(str(tmp_path), g, med, spos, rpos, src, rec, dm)
    print(f"cabi_driver {precision}", flush=True)
    r = subprocess.run([exe, str(tmp_path), precision], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rd = lambda n: np.fromfile(os.path.join(str(tmp_path), n), np.float64)  # noqa: E731
    errs = {
        "nl_fwd": rel(_split(rd("out_nl_fwd.bin"), rpos, g.nts), exp["nl_fwd"]),
        "nl_adj": rel(_split(rd("out_nl_adj.bin"), spos, g.nts), exp["nl_adj"]),
        "born_fwd": rel(_split(rd("out_born_fwd.bin"), rpos, g.nts), exp["born_fwd"]),
        "born_adj": rel(rd("out_born_adj.bin").reshape(exp["born_adj"].shape), exp["born_adj"]),
    }
    print(f"C++ caller {precision}: " + ", ".join(f"{k} {v:.2e}" for k, v in errs.items()), flush=True)
    assert all(v < tol for v in errs.values()), errs
