"""The grid file and the restart of finflo_amd.gridio: a grid written here is read back
node for node, level 2 takes every second node as the multigrid levels of the synthetic
cases do, and a Fortran program compiled with amdflang that performs the reference's own
READ sequence for the grid file (NBLOCK; IX, IY, IZ of every block; X, Y, Z per block,
main.f:1368-1374, messp.f:1127) reads the same numbers; a restart round-trips the solved
variables."""
import os
import subprocess

import numpy as np
import pytest

from finflo_amd.cases import make_hierarchy, node_grid, stretched
from finflo_amd.gridio import read_grid, read_restart, write_grid, write_restart

FC = "/opt/rocm/bin/amdflang"


def interior_nodes(nodes):
    return np.ascontiguousarray(nodes[2:-2, 2:-2, 2:-2])


def grids():
    a = node_grid(stretched(8, 2.0), stretched(6, 1.0, ratio=1.1), stretched(4, 1.0),
                  distort=0.1, seed=1)
    b = node_grid(stretched(4, 1.0), stretched(8, 1.0), stretched(2, 0.5), distort=0.05, seed=2)
    return [interior_nodes(a), interior_nodes(b)]


def test_grid_round_trip_and_levels(tmp_path):
    g = grids()
    path = tmp_path / "grid.bin"
    write_grid(path, g)
    back = read_grid(path)
    for x, y in zip(g, back):
        np.testing.assert_array_equal(x, y)
    lv2 = read_grid(path, level=2)
    np.testing.assert_array_equal(lv2[0], g[0][::2, ::2, ::2])
    # a direction of three nodes keeps every node
    np.testing.assert_array_equal(lv2[1], g[1][:, ::2, ::2])


@pytest.mark.skipif(not os.path.exists(FC), reason="amdflang not installed")
def test_grid_file_reads_in_fortran(tmp_path):
    g = grids()
    write_grid(tmp_path / "grid.bin", g)
    prog = """      PROGRAM RDGRID
      IMPLICIT NONE
      INTEGER :: NBLOCK, N, I, NGRID
      INTEGER :: IX(10), IY(10), IZ(10)
      REAL(8), ALLOCATABLE :: X(:), Y(:), Z(:)
      OPEN(41,FILE='grid.bin',STATUS='OLD',FORM='UNFORMATTED')
      READ (41) NBLOCK
      READ (41) (IX(N),IY(N),IZ(N),N=1,NBLOCK)
      DO N = 1,NBLOCK
         NGRID = IX(N)*IY(N)*IZ(N)
         ALLOCATE(X(NGRID),Y(NGRID),Z(NGRID))
         READ(41)(X(I),I=1,NGRID),(Y(I),I=1,NGRID),(Z(I),I=1,NGRID)
         WRITE(*,'(3I6,3ES25.16)') IX(N),IY(N),IZ(N),SUM(X),X(NGRID),
     +     Z(2)
         DEALLOCATE(X,Y,Z)
      ENDDO
      END PROGRAM
"""
    (tmp_path / "rdgrid.f").write_text(prog)
    subprocess.check_call([FC, "-O0", "rdgrid.f", "-o", "rdgrid"], cwd=tmp_path)
    out = subprocess.check_output(["./rdgrid"], cwd=tmp_path, text=True).split("\n")
    for line, x in zip(out, g):
        v = line.split()
        nk, nj, ni, _ = x.shape
        assert [int(t) for t in v[:3]] == [ni, nj, nk]
        flat = x[..., 0].ravel()
        assert float(v[3]) == pytest.approx(flat.sum(), rel=1e-14)
        assert float(v[4]) == flat[-1]
        assert float(v[5]) == x[..., 2].ravel()[1]


def test_restart_round_trip(tmp_path):
    levels = [make_hierarchy(8, 6, 4, 2, kind="all", amp=0.02, seed=s) for s in (1, 2)]
    path = tmp_path / "restart.bin"
    write_restart(path, levels)
    fresh = [make_hierarchy(8, 6, 4, 2, kind="all", amp=0.02, seed=9) for _ in (1, 2)]
    read_restart(path, fresh)
    for la, lb in zip(levels, fresh):
        for a, b in zip(la, lb):
            for n in ("PDIFF", "TEMP", "U", "V", "W", "RK", "REPS"):
                np.testing.assert_array_equal(a[n], b[n])
