"""The BC file of finflo_amd.bcfile (READBC gluebm.f:2553-3340; gluebm.f does not build here,
it needs an MPI module the image lacks):
* a hand-written file in the reference's format (tests/golden/bc_two_blocks.bc) against the
  ICON/RCON values READBC's rules give for it, worked out by hand: leading blanks, force-
  group letters of both cases, a face of two patches with their ranges, whole faces whose
  ranges are replaced, the wall lines with 13 (WTRAN defaults to 1) and 14 numbers, an
  inlet of type 14 (no data-file line) and an outlet of type 3 (a data-file line), the
  connection pair completed with the partner's ranges and block-local numbers, a mirror;
* the writer and reader together with the conversion to the library's registrations:
  the blocks of a multi-block case and the halves of a block in four orientations come
  back as the same boundary patches and the same one-to-one connections, orientation
  included (found from the cell centres, as PATCHE finds it from the grid);
* the number reader INLINE/UNWRAP and FGIDS on their own."""
import os

import numpy as np
import pytest

from finflo_amd.bcfile import (bc_patches, con_patches, fgids, inline, read_bc_file,
                               wall_ranges, write_bc_file, _a80)

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "bc_two_blocks.bc")


def icon(rec, *ns):
    return [rec[n] for n in ns]


def test_golden_file_gives_readbc_values():
    recs = read_bc_file(open(GOLDEN).read(), [(4, 3, 2), (4, 3, 2)], mgrid=[3, 2])
    assert len(recs) == 13
    assert [r[1] for r in recs] == [3, 8, 8, 2, 1, 7, 2, 1, 8, 4, 5, 2, 2]
    assert [r[2] for r in recs] == [1] * 7 + [2] * 6
    assert [r[3] for r in recs] == [1, 2, 2, 3, 4, 5, 6, 1, 2, 3, 4, 5, 6]
    # ranges: given on the two-patch face, the whole face elsewhere
    assert [icon(r, 4, 5, 6, 7) for r in recs] == [
        [1, 3, 1, 2], [1, 2, 1, 2], [3, 4, 1, 2], [1, 4, 1, 3], [1, 3, 1, 2], [1, 4, 1, 2],
        [1, 4, 1, 3], [1, 3, 1, 2], [1, 4, 1, 2], [1, 4, 1, 3], [1, 3, 1, 2], [1, 4, 1, 2],
        [1, 4, 1, 3]]
    assert [r[21] for r in recs] == [1, 1, 1, 1, 4, 3, 2, 1, 1, 1, 4, 3, 2]
    inl, s1, s2, out = recs[0], recs[1], recs[2], recs[10]
    assert icon(inl, 8, 27) == [14, 14] and inl.bnfile == "" and inl.groups[0] == "A"
    assert icon(out, 8, 27) == [3, 3] and out.bnfile == "outlet.dat" and out.groups[51] == "z"
    assert s1[20] == 1 and s1.rcon[4] == 295.5 and s1.rcon[0] == 300.0
    assert s1.rcon[12] == 1e-6 and s1.rcon[13] == 1.0 and s1.groups[1] == "B"
    assert s2[20] == 2 and s2.rcon[5] == 1500.0 and s2.rcon[13] == 0.0
    assert s2.groups[1] == "B" and s2.groups[28] == "c"
    # 14 numbers on the line: ICON(20) and RCON(1..13), WTRAN = 1
    assert recs[8][20] == 0 and not recs[8].rcon[:13].any() and recs[8].rcon[13] == 1.0
    # the connection pair
    c1, c2 = recs[4], recs[7]
    assert icon(c1, 8, 9, 10, 11, 12, 13) == [2, 1, 1, 3, 1, 2]
    assert icon(c2, 8, 9, 10, 11, 12, 13) == [1, 4, 1, 3, 1, 2]
    assert icon(c1, 14, 15, 16, 17, 18, 19, 28) == [0, 1, 8, 1, 1, 2, 2]
    assert icon(c2, 14, 15, 16, 17, 18, 19, 28) == [0, 5, 5, 1, 1, 3, 1]
    # the mirror
    m = recs[9]
    assert icon(m, 8, 9, 10, 11, 12, 13, 14, 15, 16, 18, 19) == [2, 3, 1, 4, 1, 3, 0, 3, 10, 1, 2]
    assert [r[24] for r in recs] == [r[2] for r in recs]
    assert [r[25] for r in recs] == list(range(1, 14))
    # the library's registrations
    bps = bc_patches(recs, 1)
    assert [(b.ibc, b.iface, b.kx1, b.kx2, b.ky1, b.ky2, b.iheat) for b in bps] == [
        (3, 1, 1, 3, 1, 2, 0), (8, 2, 1, 2, 1, 2, 1), (8, 2, 3, 4, 1, 2, 2),
        (2, 3, 1, 4, 1, 3, 0), (7, 5, 1, 4, 1, 2, 0), (2, 6, 1, 4, 1, 3, 0)]
    assert bps[0].inlout == 14 and bps[1].tsurf == 295.5
    assert wall_ranges(recs, 1) == [(2, 1, 2, 1, 2, 1), (2, 3, 4, 1, 2, 2)]


def test_number_reader():
    assert inline(_a80("  12  96 1 48   3 4 1"), 8) == [12, 96, 1, 48, 3, 4, 1]
    assert inline(_a80(" x 2.5 -1.5 7"), 2) == [3, -2]
    line, groups = fgids(_a80(" 1 2 Ab z"))
    assert groups[0] == "A" and groups[27] == "b" and groups[51] == "z"
    assert inline(line, 9) == [1, 2]


def _rt(dims, patches, blocks):
    text = write_bc_file(dims, patches)
    return read_bc_file(text, dims)


@pytest.mark.parametrize("mode", ["same", "rot180", "rot90", "swap"])
def test_connections_round_trip_with_orientation(mode):
    from tests.test_gpu_multiblock import global_block, halves
    prm, G, nodes = global_block()
    A, B, ca, cb = halves(G, nodes, mode)
    dims = [(A.imax, A.jmax, A.kmax), (B.imax, B.jmax, B.kmax)]
    pats = []
    for blk, cons, n in ((A, ca, 1), (B, cb, 2)):
        ext = {1: (blk.jmax, blk.kmax), 2: (blk.imax, blk.kmax), 3: (blk.imax, blk.jmax)}
        lst = []
        for face in range(1, 7):
            c = [t for t in cons if t[0] == face]
            na, nb = ext[(face - 1) % 3 + 1]
            if c:
                t = c[0]
                lst.append(dict(tbc="CON", face=face, ranges=t[1:5], partner=(t[5], t[6], 1)))
            else:
                lst.append(dict(tbc="EXT", face=face, ranges=(1, na, 1, nb)))
        pats.append(lst)
    recs = _rt(dims, pats, [A, B])
    assert con_patches(recs, 1, [A, B]) == ca
    assert con_patches(recs, 2, [A, B]) == cb


def test_multiblock_case_round_trip():
    from finflo_amd.cases import bump_blocks, flow_params, freestream_turbulence, gaussian_bump
    prm = flow_params(iturb=6, mach=0.5, reynolds=1e6, cfl=5.0)
    turb = freestream_turbulence(prm, limits=True)
    segs = [np.array([8, 12]), np.array([8, 4]), np.array([4, 8])]
    blocks, cons, walls, bcs = bump_blocks(prm, segs, gaussian_bump(0.05, 0.2), 1, 3,
                                           first=1e-3, turb=turb)
    fine = [lv[0] for lv in blocks]
    dims = [(b.imax, b.jmax, b.kmax) for b in fine]
    pats = []
    for n, b in enumerate(fine):
        lst = []
        for t in cons[n]:
            lst.append(dict(tbc="CON", face=t[0], ranges=t[1:5], partner=(t[5], t[6], 1)))
        for bp in bcs[n][0]:
            tbc = {2: "EXT", 8: "SOL"}[bp.ibc]
            lst.append(dict(tbc=tbc, face=bp.iface, ranges=(bp.kx1, bp.kx2, bp.ky1, bp.ky2),
                            iheat=bp.iheat, rcon=[0.0] * 14))
        pats.append(lst)
    recs = _rt(dims, pats, fine)
    for n, b in enumerate(fine):
        assert con_patches(recs, n + 1, fine) == sorted(cons[n], key=lambda t: t[0])
        got = [(p.ibc, p.iface, p.kx1, p.kx2, p.ky1, p.ky2) for p in bc_patches(recs, n + 1)]
        want = sorted((p.ibc, p.iface, p.kx1, p.kx2, p.ky1, p.ky2) for p in bcs[n][0])
        assert sorted(got) == want
