"""Write tests/golden/test3d_crop.npz from the reference's Test_3D files
(test infrastructure; run where the reference tree exists):

    python tests/golden/make_test3d_crop.py

Contents (data only):
  prim          float32 [9, NZ, NY, 192]: result_prim_0..8.000000 of
                TEST/Test_3D/ini (rho, v0, v1, v2, e_int, B0, B1, B2 and
                Spitzer's heat flux sflx; the 192-cell vertical direction
                fastest), cropped to the first NY x NZ horizontal cells
  gsize, dx     the full Test_3D grid (ini/Header.000000)
  header_ini    text of ini/Header.000000
  backup_ini    text of ini/backup.dat
  header_10     text of 3D/Header.000010, written by the reference after 10 steps
  backup_20     text of TEST/Test_3D/backup.dat, written after 20 steps
  run_10        (time, dt) at iteration 10 as anls.log prints them
  va_10         1/sqrt(inv_va2max) at iteration 10 as Header.000010 prints it
  bc_20         (iteration, time, p_bc, s_bc) of iteration 20 as BC.log and
                backup.dat print them
"""
from __future__ import annotations

import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/TEST/Test_3D"
NY = NZ = 10


def text(path):
    with open(path) as f:
        return f.read()


def main():
    hdr = text(os.path.join(SRC, "ini", "Header.000000")).split()
    gsize = np.array([int(v) for v in hdr[:3]], dtype=np.int64)
    dx = np.array([float(v) for v in hdr[3:6]], dtype=np.float64)
    n0, n1, n2 = (int(v) for v in gsize)
    prim = np.empty((9, NZ, NY, n0), dtype=np.float32)
    for v in range(9):
        a = np.fromfile(os.path.join(SRC, "ini", f"result_prim_{v}.000000"), dtype=np.float32)
        prim[v] = a.reshape(n2, n1, n0)[:NZ, :NY, :]

    anls = text(os.path.join(SRC, "anls.log")).split("\n")[0].split()
    assert anls[0] == "10"
    run_10 = np.array([float(anls[1]), float(anls[2])])
    header_10 = text(os.path.join(SRC, "3D", "Header.000010"))
    va_10 = float(header_10.split()[8])
    backup_20 = text(os.path.join(SRC, "backup.dat"))
    b = backup_20.split()
    bc_20 = np.array([float(b[0]), float(b[1]), float(b[2]), float(b[3])])

    np.savez_compressed(
        os.path.join(HERE, "test3d_crop.npz"),
        prim=prim, gsize=gsize, dx=dx,
        header_ini=np.array(text(os.path.join(SRC, "ini", "Header.000000"))),
        backup_ini=np.array(text(os.path.join(SRC, "ini", "backup.dat"))),
        header_10=np.array(header_10), backup_20=np.array(backup_20),
        run_10=run_10, va_10=np.array(va_10), bc_20=bc_20)


if __name__ == "__main__":
    main()
