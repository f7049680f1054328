"""Multi-block on the HIP path: the connection halo exchange (ff_con_register,
ff_halo_exchange) and the multi-block cycle (ff_vcycle_blocks), through the C-ABI.
A block is cut in two, the second half either in the same index orientation or rotated
(its I axis along the global K, its K axis against the global I); the truth is the uncut
block:
* after the geometry exchange the ghost cells behind the cut hold the neighbour's cell
  centres and volumes (bit for bit in the same orientation, to rounding when rotated,
  whose sub-block geometry sums its nodes in another order);
* after the flow exchange the residual of both halves is the uncut block's residual
  (bit for bit in the same orientation, 1e-12 when rotated: the directions of the
  face sweeps differ there);
* multi-block V-cycles against the oracle's V-cycle of each block with the exchange done
  on the host from the global cell numbering."""
import numpy as np
import pytest

from finflo_amd.cases import fill_geometry, fill_state, node_grid, stretched, sub_block
from finflo_amd.layout import Block, default_params
from oracle.oracle_py import OracleHierarchy

pytestmark = pytest.mark.gpu

RES = ["DRO", "DM", "DN", "DW", "DE"]


def global_block(iturb=0, dims=(24, 16, 12), seed=3):
    prm = default_params(iturb=iturb, cfll=3.0)
    nodes = node_grid(stretched(dims[0], 2.0), stretched(dims[1], 1.0, ratio=1.1),
                      stretched(dims[2], 1.0), distort=0.1, seed=seed)
    G = Block(*dims)
    fill_geometry(G, nodes)
    fill_state(G, prm, seed=seed, amp=0.03, kind="all", turb=(1.0, 80.0) if iturb else None)
    return prm, G, nodes


def halves(G, nodes, mode):
    """The two halves of G cut at I = IMAX/2, the second in orientation `mode`: "same";
    "rot180" (its I and J against the global ones); "rot90" (its I along the global J, its J
    against the global I); "swap" (its I along the global K, its K against the global I,
    which turns the in-plane axes of the cut face)."""
    h = G.imax // 2
    A = sub_block(G, nodes, (1, 1, 1), (h, G.jmax, G.kmax))
    lo, hi = (h + 1, 1, 1), (G.imax, G.jmax, G.kmax)
    if mode == "same":
        B = sub_block(G, nodes, lo, hi)
        ca = [(4, 1, G.jmax, 1, G.kmax, 2, 1, 1, 1, 0, 1, 1)]
        cb = [(1, 1, G.jmax, 1, G.kmax, 1, 4, 1, 1, 0, 1, 1)]
    elif mode == "rot180":
        B = sub_block(G, nodes, lo, hi, perm=(0, 1, 2), flip=(True, True, False))
        ca = [(4, 1, G.jmax, 1, G.kmax, 2, 4, G.jmax, 1, 0, -1, 1)]
        cb = [(4, 1, G.jmax, 1, G.kmax, 1, 4, G.jmax, 1, 0, -1, 1)]
    elif mode == "rot90":
        B = sub_block(G, nodes, lo, hi, perm=(1, 0, 2), flip=(False, True, False))
        ca = [(4, 1, G.jmax, 1, G.kmax, 2, 5, 1, 1, 0, 1, 1)]
        cb = [(5, 1, G.jmax, 1, G.kmax, 1, 4, 1, 1, 0, 1, 1)]
    else:
        B = sub_block(G, nodes, lo, hi, perm=(2, 1, 0), flip=(False, False, True))
        ca = [(4, 1, G.jmax, 1, G.kmax, 2, 6, 1, 1, 1, 1, 1)]
        cb = [(6, 1, G.kmax, 1, G.jmax, 1, 4, 1, 1, 1, 1, 1)]
    return A, B, ca, cb


def to_global_b(G, Bfield, mode):
    """A (KMAX, JMAX, IMAX) interior array of B in the global (k, j, i) order of its cells."""
    if mode == "same":
        return Bfield
    if mode == "rot180":
        return Bfield[:, ::-1, ::-1]
    if mode == "rot90":
        return np.transpose(Bfield, (0, 2, 1))[:, :, ::-1]
    t = np.transpose(Bfield, (2, 1, 0))        # [i'][j'][k'] = [k][j][reversed i]
    return np.flip(t, axis=2)


def context(prm, blocks, cons):
    from finflo_amd.ffhip import Context
    ctx = Context(prm)
    for n, b in enumerate(blocks, start=1):
        ctx.register(b, block=n)
        ctx.upload(block=n)
    for n, cp in enumerate(cons, start=1):
        ctx.con_register(cp, block=n)
    return ctx


@pytest.mark.parametrize("mode", ["same", "swap"])
def test_geometry_exchange(hip_device, mode):
    print(f"case connection geometry {mode}", flush=True)
    rotated = mode != "same"
    prm, G, nodes = global_block()
    A, B, ca, cb = halves(G, nodes, mode)
    ctx = context(prm, [A, B], [ca, cb])
    ctx.halo_exchange(what="geometry")
    ctx.download(block=1)
    ctx.download(block=2)
    h = G.imax // 2
    tol = 0.0 if not rotated else 1e-13
    for f in ("XC", "YC", "ZC", "VOL"):
        ga = A.view3(f)[2:-2, 2:-2, h + 2:h + 4]            # A's ghosts I = h+1, h+2
        want = G.view3(f)[2:-2, 2:-2, h + 2:h + 4]
        err = np.abs(ga - want).max() / np.abs(want).max()
        assert err <= tol, (f, err)
        gb = B.view3(f)
        if not rotated:
            got = gb[2:-2, 2:-2, 0:2]                       # B's ghosts I = -1, 0
        else:
            # B's ghosts K' = KMAX+2, KMAX+1 are the global I = h-1, h
            got = np.transpose(np.flip(gb[-2:, 2:-2, 2:-2], axis=0), (2, 1, 0))
        want = G.view3(f)[2:-2, 2:-2, h:h + 2]
        err = np.abs(got - want).max() / np.abs(want).max()
        assert err <= tol, (f, "B", err)
    ctx.close()


@pytest.mark.parametrize("mode", ["same", "rot180", "rot90"])
def test_split_residual_is_the_uncut_residual(hip_device, mode):
    """The K axis is kept: the reference's FLUXKE takes DERIV's one-sided branch on every
    K-face (base3c.f:719), so a block whose K axis runs along another global direction has,
    as in the reference, a different residual there."""
    from finflo_amd.ffhip import Context
    print(f"case split residual {mode}", flush=True)
    prm, G, nodes = global_block()
    A, B, ca, cb = halves(G, nodes, mode)
    ctx = context(prm, [A, B], [ca, cb])
    ctx.halo_exchange(what="geometry")
    ctx.halo_exchange(what="flow")
    for n in (1, 2):
        ctx.statim(block=n)
        ctx.turbdi(block=n)
        ctx.residi(block=n)
        ctx.download(block=n)
    cg = Context(prm)
    cg.register(G)
    cg.upload()
    cg.statim()
    cg.turbdi()
    cg.residi()
    cg.download()
    h = G.imax // 2
    for f in RES:
        want = G.interior(f)
        scale = np.abs(want).max()
        got_a = A.interior(f)
        got_b = to_global_b(G, B.interior(f), mode)
        if mode == "same":
            np.testing.assert_array_equal(got_a, want[:, :, :h], err_msg=f)
            np.testing.assert_array_equal(got_b, want[:, :, h:], err_msg=f)
        else:
            # the I and J sweeps of the rotated half run in another order: rounding only
            assert np.abs(got_a - want[:, :, :h]).max() <= 1e-12 * scale, f
            assert np.abs(got_b - want[:, :, h:]).max() <= 1e-12 * scale, f
    ctx.close()
    cg.close()


def test_multiblock_vcycles(hip_device):
    """Two halves, 2 multigrid levels each, 3 multi-block cycles against the oracle's
    V-cycles of each half with the level-1 exchange done on the host."""
    from finflo_amd.cases import coarsen_nodes
    from finflo_amd.ffhip import Context
    print("case multi-block V-cycles", flush=True)
    prm, G, nodes = global_block(dims=(24, 16, 12))
    h = G.imax // 2
    levels = []
    for lo, hi_ in (((1, 1, 1), (h, G.jmax, G.kmax)), ((h + 1, 1, 1), (G.imax, G.jmax, G.kmax))):
        fine = sub_block(G, nodes, lo, hi_)
        # the coarse level of each half from its own fine nodes
        g = 2
        sub = nodes[g + lo[2] - 1:g + hi_[2] + 1, g + lo[1] - 1:g + hi_[1] + 1, g + lo[0] - 1:g + hi_[0] + 1]
        from finflo_amd.cases import _extend_axis
        full = sub
        for ax in range(3):
            full = _extend_axis(full, ax)
        cn = coarsen_nodes(full)
        coarse = Block(fine.imax // 2, fine.jmax // 2, fine.kmax // 2)
        fill_geometry(coarse, cn)
        fill_state(coarse, prm, seed=0, amp=0.0, kind="density")
        levels.append([fine, coarse])
    host = [[b.copy() for b in lv] for lv in levels]
    ctx = Context(prm)
    for n, lv in enumerate(levels, start=1):
        for m, b in enumerate(lv, start=1):
            ctx.register(b, level=m, block=n)
            ctx.upload(level=m, block=n)
    ctx.con_register([(4, 1, G.jmax, 1, G.kmax, 2, 1, 1, 1, 0, 1, 1)], block=1)
    ctx.con_register([(1, 1, G.jmax, 1, G.kmax, 1, 4, 1, 1, 0, 1, 1)], block=2)
    ctx.halo_exchange(what="geometry")
    ohs = [OracleHierarchy(prm, lv) for lv in host]
    # the host exchange: A's ghosts I = h+1, h+2 <- B's cells 1, 2 and B's I = 0, -1 <- A's h, h-1
    exchange_host(host[0][0], host[1][0], geometry=True)
    for _ in range(3):
        r = np.zeros(7)
        for oh in ohs:
            r = r + oh.vcycle()
        exchange_host(host[0][0], host[1][0])
    got = ctx.vcycle_blocks([1, 2], nlev=2, ncycles=3)
    for n in (1, 2):
        for m in (1, 2):
            ctx.download(level=m, block=n)
    for n in range(2):
        for m in range(2):
            for f in ("TEMP", "U", "V", "W", "PDIFF"):
                np.testing.assert_array_equal(levels[n][m][f], host[n][m][f],
                                              err_msg=f"block {n + 1} level {m + 1} {f}")
    np.testing.assert_allclose(got[:5], r[:5], rtol=1e-12)
    ctx.close()


FLOW = ["PDIFF", "TEMP", "U", "V", "W", "RK", "REPS", "P", "RO", "E", "C", "VIS", "CH", "CP",
        "DRDP", "DRDH", "RM", "RN", "RW", "EPS2", "VIST"]
GEOM = ["VOL", "XC", "YC", "ZC", "DISTW", "D1", "D2", "D3"]


def exchange_host(A, B, geometry=False):
    """The I-cut exchange of two same-orientation halves on the host (test infrastructure)."""
    ia = A.imax
    for f in (GEOM if geometry else FLOW):
        va, vb = A.view3(f), B.view3(f)
        va[2:-2, 2:-2, ia + 2:ia + 4] = vb[2:-2, 2:-2, 2:4]
        vb[2:-2, 2:-2, 0:2] = va[2:-2, 2:-2, ia:ia + 2]


def exchange_con_host(blocks, cons, fields):
    """The level-1 exchange of same-orientation full-face connections on the host (test
    infrastructure): for each connection tuple (iface, ka1, ka2, kb1, kb2, nblock, nface, 1,
    1, 0, 1, 1) of block n, its two ghost layers take the connected block's two layers next
    to the connected face."""
    def layer(b, face, depth):
        """(K, J, I)-view slice of the cell layer `depth` in from `face` (0, -1: ghosts),
        in-plane ranges 1..max."""
        d = (face - 1) % 3
        nd = (b.imax, b.jmax, b.kmax)[d]
        idx = depth if face <= 3 else nd + 1 - depth
        sl = [slice(2, -2)] * 3                  # view3 order (k, j, i)
        sl[2 - d] = slice(idx + 1, idx + 2)
        return tuple(sl)
    for n, cl in enumerate(cons):
        A = blocks[n]
        for (iface, ka1, ka2, kb1, kb2, nb, nface, *_rest) in cl:
            B = blocks[nb - 1]
            for f in fields:
                va, vb = A.view3(f), B.view3(f)
                for l in (1, 2):
                    va[layer(A, iface, 1 - l)] = vb[layer(B, nface, l)]


def test_config3_multiblock_cycles(hip_device):
    """The 24-block bump case at full size (about 7.6M cells, 3 levels, SST, walls on the
    bottom row): 2 multi-block cycles against the oracle's V-cycles of every block with the
    exchange done on the host, at 1e-12 of each field's largest magnitude; then the captured
    multi-block graph against the eager cycles of a twin context, bit for bit."""
    from finflo_amd.cases import config3
    from finflo_amd.ffhip import Context
    print("case config 3: 24 blocks, 2 multi-block cycles", flush=True)
    prm, blocks, cons, walls, bcs = config3(seed=3, nlev=3)
    nb = len(blocks)
    host = [[b.copy() for b in lv] for lv in blocks]
    twin = [[b.copy() for b in lv] for lv in blocks]
    ctxs = []
    for lvs in (blocks, twin):
        ctx = Context(prm)
        for n, lv in enumerate(lvs, start=1):
            for m, b in enumerate(lv, start=1):
                ctx.register(b, level=m, block=n)
                ctx.upload(level=m, block=n)
                if walls[n - 1][m - 1] is not None:
                    ctx.wall_register(*walls[n - 1][m - 1], level=m, block=n)
                ctx.bc_register(bcs[n - 1][m - 1], level=m, block=n)
        for n in range(1, nb + 1):
            ctx.con_register(cons[n - 1], block=n)
        for n in range(1, nb + 1):
            for m in range(1, 4):
                ctx.ghostext(level=m, block=n)
        ctx.halo_exchange(what="geometry")
        ctx.halo_exchange(what="flow")
        ctxs.append(ctx)
    ctx, ctx2 = ctxs
    # the same start on the host: GHOSTEXT, the geometry and flow exchanges
    ohs = []
    for n, lv in enumerate(host):
        oh = OracleHierarchy(prm, lv)
        for m in range(3):
            if walls[n][m] is not None:
                wp, ps = walls[n][m]
                oh.levels[m].set_walls(wp, [p.copy() for p in ps])
            oh.levels[m].set_bcs(bcs[n][m])
            oh.levels[m].ghostext()
        ohs.append(oh)
        # only the last block's sets stay registered: each block activates its own later
    fine = [lv[0] for lv in host]
    exchange_con_host(fine, cons, GEOM)
    exchange_con_host(fine, cons, FLOW)
    for _ in range(2):
        r = np.zeros(7)
        for n, oh in enumerate(ohs):
            # the oracle keeps walls and patches per level: this block's before its cycle
            for m in range(3):
                oh.levels[m].activate()
            r = r + oh.vcycle()
        exchange_con_host(fine, cons, FLOW)
    got = ctx.vcycle_blocks(list(range(1, nb + 1)), nlev=3, ncycles=2)
    for n in range(nb):
        ctx.download(level=1, block=n + 1)
        for f in ("TEMP", "U", "V", "W", "PDIFF", "RK", "REPS"):
            x, y = blocks[n][0][f], host[n][0][f]
            assert np.abs(x - y).max() <= 1e-12 * np.abs(y).max(), (n + 1, f)
    np.testing.assert_allclose(got[:7], r[:7], rtol=1e-9)
    # the graph: two replays against two eager cycles of the twin
    ctx2.vcycle_blocks(list(range(1, nb + 1)), nlev=3, ncycles=2)
    ctx.vcycle_blocks_graph(list(range(1, nb + 1)), nlev=3, ncycles=2)
    ctx2.vcycle_blocks(list(range(1, nb + 1)), nlev=3, ncycles=2)
    for n in range(nb):
        for m in range(1, 4):
            ctx.download(level=m, block=n + 1)
            ctx2.download(level=m, block=n + 1)
            np.testing.assert_array_equal(blocks[n][m - 1].F, twin[n][m - 1].F,
                                          err_msg=f"graph block {n + 1} level {m}")
    # the blocks side by side on the stream pool against one after the other, bit for bit,
    # residual sums included
    print("case config 3: blocks on the stream pool", flush=True)
    ctx.block_streams(True)
    r1 = ctx.vcycle_blocks(list(range(1, nb + 1)), nlev=3, ncycles=2)
    r2 = ctx2.vcycle_blocks(list(range(1, nb + 1)), nlev=3, ncycles=2)
    np.testing.assert_array_equal(r1, r2)
    for n in range(nb):
        for m in range(1, 4):
            ctx.download(level=m, block=n + 1)
            ctx2.download(level=m, block=n + 1)
            np.testing.assert_array_equal(blocks[n][m - 1].F, twin[n][m - 1].F,
                                          err_msg=f"stream pool block {n + 1} level {m}")
    ctx.close()
    ctx2.close()


def test_graph_capture_limit(hip_device):
    """ff_vcycle_blocks_graph refuses, before touching any block, a capture of more than 3
    blocks x 3 levels per branch of its eight-stream pool (FF_E_UNSUP)."""
    from finflo_amd.ffhip import Context, FFError
    from finflo_amd.layout import default_params
    print("case multi-block graph capture limit", flush=True)
    ctx = Context(default_params())
    with pytest.raises(FFError):
        ctx.vcycle_blocks_graph(list(range(1, 26)), 3, ncycles=1)
    with pytest.raises(FFError):
        ctx.vcycle_blocks_graph(list(range(1, 9)), 10, ncycles=1)
    ctx.close()
