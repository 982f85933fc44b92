"""Model padding tool (the reference's padElasticFileGpuMain_3D.cpp:7-193).

    python -m elasticisolib_3d_amd.pad model=in.H data=out.H zPad=30 xPad=30 yPad=30 \
        [fat=4] [blockSize=16] [freeSurface=0]

Pads every parameter of a [nz, nx, ny(, nPar)] SEP model by edge replication:
zPad / xPad cells before the model, zPadPlus / xPadPlus after it so that the
padded interior (model + pads) is a multiple of blockSize, yPad on both y
sides, plus FAT cells on every face; origins move by -(fat + pad) * d. With
freeSurface=1 the top pad is FAT cells (the free surface sits at the top).
Prints the padding parameters that the propagation par file needs.
"""
from __future__ import annotations

import math
import sys
from typing import Dict

import numpy as np

from .sep import Axis, Hypercube, ParObject, Vector, read_vector, write_vector


def padding(nz: int, nx: int, ny: int, z_pad: int, x_pad: int, y_pad: int, fat: int = 4, block: int = 16,
            free_surface: int = 0) -> Dict[str, int]:
    """The padding rule of padElasticFileGpuMain_3D.cpp:47-86."""
    if free_surface == 0:
        nb_z = math.ceil(float(np.float32(z_pad * 2 + nz) / np.float32(block)))
        z_minus = z_pad
    elif free_surface == 1:
        nb_z = math.ceil(float(np.float32(z_pad + nz + fat) / np.float32(block)))
        z_minus = fat
    else:
        raise ValueError("ERROR UNKNOWN SURFACE CONDITION PARAMETER")
    z_plus = nb_z * block - nz - z_minus
    nb_x = math.ceil(float(np.float32(x_pad * 2 + nx) / np.float32(block)))
    x_plus = nb_x * block - nx - x_pad
    return {"zPadMinus": z_minus, "zPadPlus": z_plus, "xPadMinus": x_pad, "xPadPlus": x_plus,
            "yPadMinus": y_pad, "yPadPlus": y_pad, "fat": fat, "blockSize": block,
            "nz": z_minus + z_plus + nz + 2 * fat, "nx": x_pad + x_plus + nx + 2 * fat, "ny": 2 * y_pad + ny + 2 * fat}


def pad_vector(model: Vector, z_pad: int, x_pad: int, y_pad: int, fat: int = 4, block: int = 16,
               free_surface: int = 0):
    """(padded Vector [nz', nx', ny', nPar], padding dict)."""
    h = model.getHyper()
    axes = list(h.axes) + [Axis(1)] * (4 - h.getNdim())
    nz, nx, ny = axes[0].n, axes[1].n, axes[2].n
    p = padding(nz, nx, ny, z_pad, x_pad, y_pad, fat, block, free_surface)
    arr = np.asarray(model.getNdArray()).reshape(axes[3].n, ny, nx, nz)
    out = np.pad(arr, [(0, 0), (fat + y_pad, fat + y_pad), (fat + x_pad, fat + p["xPadPlus"]),
                       (fat + p["zPadMinus"], fat + p["zPadPlus"])], mode="edge")
    za, xa, ya = axes[0], axes[1], axes[2]
    hyp = Hypercube([Axis(p["nz"], za.o - (fat + p["zPadMinus"]) * za.d, za.d),
                     Axis(p["nx"], xa.o - (fat + x_pad) * xa.d, xa.d),
                     Axis(p["ny"], ya.o - (fat + y_pad) * ya.d, ya.d), Axis(axes[3].n, axes[3].o, axes[3].d)])
    return Vector(hyp, out.astype(model.getNdArray().dtype)), p


def main(argv) -> int:
    par = ParObject(argv)
    model = read_vector(par.getString("model"))
    padded, p = pad_vector(model, par.getInt("zPad"), par.getInt("xPad"), par.getInt("yPad"), par.getInt("fat", 4),
                           par.getInt("blockSize", 16), par.getInt("freeSurface", 0))
    write_vector(par.getString("data"), padded)
    for k in ("zPadMinus", "zPadPlus", "xPadMinus", "xPadPlus", "yPadMinus", "yPadPlus", "blockSize", "fat"):
        print(f"{k} = {p[k]} [samples]")
    print(f"New nz = {p['nz']} [samples including padding and FAT]")
    print(f"New nx = {p['nx']} [samples including padding and FAT]")
    print(f"New ny = {p['ny']} [samples including padding and FAT]")
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
