"""ctypes binding of libffhip.so, mirroring FINFLO's per-block call surface.

The method names follow the reference's dispatchers called from CYCLES (ns3c.f:1354-2349):
``statim`` (STATIM), ``turbdi`` (TURBDI), ``residi`` (RESIDI), ``impsdi`` (IMPSDI) and ``fin``
(FIN + limiters).  The library must be built (finflo_amd/build.py); a missing or unloadable
library raises — there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

from .layout import (FID, Block, BlockView, Params, WallParams, bc_views, con_views,
                     wall_views)

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libffhip.so")

EXPORTS = [
    "ff_ctx_create", "ff_ctx_destroy", "ff_ctx_set_stream", "ff_ctx_set_option",
    "ff_last_error", "ff_ctx_profile",
    "ff_kernel_stats", "ff_kernel_stats_reset", "ff_block_register",
    "ff_block_register_device", "ff_sync", "ff_finish", "ff_state", "ff_turb", "ff_residual",
    "ff_implicit", "ff_update", "ff_save_old", "ff_cycle", "ff_restrict", "ff_forcing",
    "ff_prolong", "ff_vcycle", "ff_vcycle_graph", "ff_wall_register", "ff_wall_sync", "ff_wall_forces",
    "ff_mlp_set", "ff_mlp_get", "ff_earsm_samples", "ff_mlp_fit_norm", "ff_mlp_infer",
    "ff_mlp_grad", "ff_mlp_train", "ff_residual_noalloc", "ff_dromax",
    "ff_bc_register", "ff_reflec", "ff_ghostext", "ff_inlout", "ff_bc_cycle", "ff_bc_read",
    "ff_con_register", "ff_halo_exchange",
    "ff_vcycle_blocks", "ff_vcycle_blocks_graph", "ff_tuned_shape",
]

FF_OPT_LUSGS, FF_LUSGS_TILES, FF_LUSGS_PLANES, FF_LUSGS_KMARCH, FF_LUSGS_STAGED = 1, 0, 1, 2, 3
FF_LUSGS_STAGED1 = 4
FF_OPT_RESIDUAL, FF_RESIDUAL_TILES, FF_RESIDUAL_CELLS = 2, 0, 1
FF_OPT_EARSM, FF_EARSM_EXACT, FF_EARSM_MLP = 3, 0, 1
FF_OPT_PRECISION, FF_PREC_FP64, FF_PREC_MIXED = 4, 0, 1
FF_OPT_AUTOTUNE, FF_OPT_TILE_SHAPE, FF_OPT_BLOCK_STREAMS = 5, 6, 7
FF_MLP_NPARAM, FF_MLP_NNORM = 4739, 26

_lib = None


class FFError(RuntimeError):
    pass


def lib():
    """Load libffhip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise FFError(f"{LIBPATH} not built: run python -m finflo_amd.build")
        L = C.CDLL(LIBPATH)
        vp, i, d = C.c_void_p, C.c_int, C.c_double
        L.ff_ctx_create.argtypes = [C.POINTER(vp), C.POINTER(Params)]
        L.ff_ctx_destroy.argtypes = [vp]
        L.ff_ctx_set_stream.argtypes = [vp, vp]
        L.ff_ctx_set_option.argtypes = [vp, i, i]
        L.ff_last_error.argtypes = [vp]
        L.ff_last_error.restype = C.c_char_p
        L.ff_ctx_profile.argtypes = [vp, i]
        L.ff_kernel_stats.argtypes = [vp, C.c_char_p, C.POINTER(C.c_longlong), C.POINTER(d)]
        L.ff_kernel_stats_reset.argtypes = [vp]
        L.ff_block_register.argtypes = [vp, i, i, C.POINTER(BlockView)]
        L.ff_block_register_device.argtypes = [vp, i, i, C.POINTER(BlockView)]
        L.ff_sync.argtypes = [vp, i, i, i, C.POINTER(C.c_int), i]
        L.ff_finish.argtypes = [vp]
        for fn in ("ff_state", "ff_turb", "ff_update", "ff_save_old"):
            getattr(L, fn).argtypes = [vp, i, i]
        L.ff_residual.argtypes = [vp, i, i, i, vp]
        L.ff_residual_noalloc.argtypes = [vp, i, C.POINTER(BlockView), i, vp]
        L.ff_dromax.argtypes = [vp, i, i, C.POINTER(d), C.POINTER(i), C.POINTER(i)]
        L.ff_bc_register.argtypes = [vp, i, i, vp, i]
        L.ff_reflec.argtypes = [vp, i, i]
        L.ff_ghostext.argtypes = [vp, i, i, d, d]
        L.ff_inlout.argtypes = [vp, i, i]
        L.ff_bc_cycle.argtypes = [vp, i, i, i]
        L.ff_bc_read.argtypes = [vp, i, i, i, vp, vp]
        L.ff_con_register.argtypes = [vp, i, i, vp, i]
        L.ff_halo_exchange.argtypes = [vp, i, i]
        L.ff_vcycle_blocks.argtypes = [vp, vp, i, i, i, vp]
        L.ff_vcycle_blocks_graph.argtypes = [vp, vp, i, i, i]
        L.ff_tuned_shape.argtypes = [vp, i, i, vp, vp]
        L.ff_implicit.argtypes = [vp, i, i, i]
        L.ff_cycle.argtypes = [vp, i, i, vp]
        L.ff_restrict.argtypes = [vp, i, i]
        L.ff_forcing.argtypes = [vp, i, i, i]
        L.ff_prolong.argtypes = [vp, i, i]
        L.ff_vcycle.argtypes = [vp, i, i, i, vp]
        L.ff_vcycle_graph.argtypes = [vp, i, i, i]
        L.ff_wall_register.argtypes = [vp, i, i, C.POINTER(WallParams), vp, i]
        L.ff_wall_sync.argtypes = [vp, i, i, i]
        L.ff_wall_forces.argtypes = [vp, i, i, vp]
        L.ff_mlp_set.argtypes = [vp, vp, vp]
        L.ff_mlp_get.argtypes = [vp, vp, vp]
        L.ff_earsm_samples.argtypes = [vp, i, i, vp, vp]
        L.ff_mlp_fit_norm.argtypes = [vp, vp, vp, i]
        L.ff_mlp_infer.argtypes = [vp, vp, i, vp]
        L.ff_mlp_grad.argtypes = [vp, vp, vp, i, vp]
        L.ff_mlp_train.argtypes = [vp, vp, vp, i, i, i, C.c_float, C.c_uint, vp]
        for fn in EXPORTS:
            if fn != "ff_last_error":
                getattr(L, fn).restype = C.c_int
        _lib = L
    return _lib


class Context:
    """One ff_ctx: parameters, a HIP stream and registered blocks."""

    def __init__(self, prm: Params):
        self.prm = prm
        self.h = C.c_void_p()
        rc = lib().ff_ctx_create(C.byref(self.h), C.byref(prm))
        if rc != 0:
            raise FFError(f"ff_ctx_create failed with code {rc}")
        self._views = {}

    def _chk(self, rc, what):
        if rc != 0:
            msg = lib().ff_last_error(self.h)
            raise FFError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def close(self):
        if self.h:
            lib().ff_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_option(self, option, value):
        self._chk(lib().ff_ctx_set_option(self.h, option, value), "ff_ctx_set_option")

    def lusgs_mode(self, mode):
        """LU-SGS sweep form: "tiles" ((J,K) tiles marching I), "kmarch" ((I,J) tiles
        marching K) or "planes" (one launch per hyperplane)."""
        self.set_option(FF_OPT_LUSGS, {"tiles": FF_LUSGS_TILES, "planes": FF_LUSGS_PLANES,
                                       "kmarch": FF_LUSGS_KMARCH,
                                       "staged": FF_LUSGS_STAGED,
                                       "staged1": FF_LUSGS_STAGED1}[mode])

    def precision(self, mode):
        """"fp64" (everything FP64) or "mixed" (the LU-SGS records in FP32)."""
        self.set_option(FF_OPT_PRECISION, {"fp64": FF_PREC_FP64, "mixed": FF_PREC_MIXED}[mode])

    def autotune(self, on=True):
        """Let the library time the residual kernels' tile shapes per block size."""
        self.set_option(FF_OPT_AUTOTUNE, 1 if on else 0)

    def block_streams(self, on=True):
        """Run the blocks of a multi-block cycle side by side on a pool of streams."""
        self.set_option(FF_OPT_BLOCK_STREAMS, 1 if on else 0)

    def tile_shape(self, s):
        """Run every residual in the autotuner's candidate shape s (-1: back to the choice)."""
        self.set_option(FF_OPT_TILE_SHAPE, s)

    def tuned_shape(self, level=1, block=1):
        """(shape (I, J, K), mean ms) the autotuner chose for a block level, or None."""
        sh = np.zeros(3, dtype=np.int32)
        ms = C.c_double()
        rc = lib().ff_tuned_shape(self.h, level, block, sh.ctypes.data, C.byref(ms))
        return (tuple(int(x) for x in sh), ms.value) if rc == 0 else None

    def residual_mode(self, mode):
        """Residual form: "tiles" (each face once) or "cells" (both faces per cell)."""
        self.set_option(FF_OPT_RESIDUAL, {"tiles": FF_RESIDUAL_TILES, "cells": FF_RESIDUAL_CELLS}[mode])

    # ---- blocks ------------------------------------------------------------------
    def register(self, blk: Block, level=1, block=1):
        v = blk.view_struct()
        self._views[(level, block)] = (blk, v)
        self._chk(lib().ff_block_register(self.h, level, block, C.byref(v)), "ff_block_register")

    @staticmethod
    def _fields(fields):
        if fields is None:
            return None, -1
        arr = (C.c_int * len(fields))(*[FID[f] for f in fields])
        return arr, len(fields)

    def upload(self, level=1, block=1, fields=None):
        arr, n = self._fields(fields)
        self._chk(lib().ff_sync(self.h, level, block, 0, arr, n), "ff_sync(h2d)")

    def download(self, level=1, block=1, fields=None):
        arr, n = self._fields(fields)
        self._chk(lib().ff_sync(self.h, level, block, 1, arr, n), "ff_sync(d2h)")

    def finish(self):
        self._chk(lib().ff_finish(self.h), "ff_finish")

    # ---- the reference's call surface ------------------------------------------------
    def statim(self, level=1, block=1):
        self._chk(lib().ff_state(self.h, level, block), "ff_state")

    def turbdi(self, level=1, block=1):
        self._chk(lib().ff_turb(self.h, level, block), "ff_turb")

    def residi(self, level=1, block=1, m2=1):
        resi = np.zeros(7)
        self._chk(lib().ff_residual(self.h, level, block, m2, resi.ctypes.data), "ff_residual")
        return resi

    def impsdi(self, level=1, block=1, m2=1):
        self._chk(lib().ff_implicit(self.h, level, block, m2), "ff_implicit")

    def fin(self, level=1, block=1):
        self._chk(lib().ff_update(self.h, level, block), "ff_update")

    def save_old(self, level=1, block=1):
        self._chk(lib().ff_save_old(self.h, level, block), "ff_save_old")

    def cycle(self, block=1, ncycles=1):
        resi = np.zeros(7)
        self._chk(lib().ff_cycle(self.h, block, ncycles, resi.ctypes.data), "ff_cycle")
        return resi

    def dromax(self, state=(1e-20, 1, 0), block=1, idrxx=6):
        """DROMXX of block N on the running state (DROMAX, IXERR, INERR) of SUBIN's block
        loop; returns the new state."""
        dm, ix, inn = C.c_double(state[0]), C.c_int(state[1]), C.c_int(state[2])
        self._chk(lib().ff_dromax(self.h, block, idrxx, C.byref(dm), C.byref(ix), C.byref(inn)),
                  "ff_dromax")
        return dm.value, ix.value, inn.value

    # ---- boundary patches (REFLEC, GHOSTEXT, INLET/OUTLET) ---------------------------------
    def bc_register(self, patches, level=1, block=1):
        """Boundary patches (finflo_amd.layout.BcPatch) of a block level; ff_state then ends
        with REFLEC over them, and a level-1 residual starts its fluxes with INLET/OUTLET
        over the INL/OUT patches."""
        views = bc_views(patches)
        self._chk(lib().ff_bc_register(self.h, level, block, views, len(patches)), "ff_bc_register")

    def reflec(self, level=1, block=1):
        self._chk(lib().ff_reflec(self.h, level, block), "ff_reflec")

    def ghostext(self, level=1, block=1, alpha=0.0, beta=0.0):
        self._chk(lib().ff_ghostext(self.h, level, block, alpha, beta), "ff_ghostext")

    def inlout(self, level=1, block=1):
        """INLET/OUTLET over the registered INL/OUT patches."""
        self._chk(lib().ff_inlout(self.h, level, block), "ff_inlout")

    def bc_cycle(self, icycle, level=1, block=1):
        """Set the block's cycle count ICYCLE."""
        self._chk(lib().ff_bc_cycle(self.h, level, block, icycle), "ff_bc_cycle")

    def bc_read(self, patch, size, level=1, block=1):
        """BOUNP and BOUNPD of INL/OUT patch `patch` as two arrays of `size` entries."""
        bp, bpd = np.zeros(size), np.zeros(size)
        self._chk(lib().ff_bc_read(self.h, level, block, patch, bp.ctypes.data, bpd.ctypes.data),
                  "ff_bc_read")
        return bp, bpd

    # ---- block connections -------------------------------------------------------------
    def con_register(self, patches, level=1, block=1):
        """Connection patches of a block level: tuples (iface, ka1, ka2, kb1, kb2, nblock,
        nface, na0, nb0, swap, da, db)."""
        self._chk(lib().ff_con_register(self.h, level, block, con_views(patches), len(patches)),
                  "ff_con_register")

    def halo_exchange(self, level=1, what="flow"):
        self._chk(lib().ff_halo_exchange(self.h, level, {"flow": 0, "geometry": 1}[what]),
                  "ff_halo_exchange")

    def vcycle_blocks(self, blocks, nlev, ncycles=1):
        """ncycles multi-block cycles over the listed blocks; the RESL2 sums of the last."""
        arr = (C.c_int * len(blocks))(*blocks)
        resi = np.zeros(7)
        self._chk(lib().ff_vcycle_blocks(self.h, arr, len(blocks), nlev, ncycles,
                                         resi.ctypes.data), "ff_vcycle_blocks")
        return resi

    def vcycle_blocks_graph(self, blocks, nlev, ncycles=1):
        """ncycles multi-block cycles replayed from one captured graph (no residual sums)."""
        arr = (C.c_int * len(blocks))(*blocks)
        self._chk(lib().ff_vcycle_blocks_graph(self.h, arr, len(blocks), nlev, ncycles),
                  "ff_vcycle_blocks_graph")

    # ---- multigrid (TRANSF/TRANKE, FORCFU/ADDRES, INTERP/INTERT, the V-cycle) ----------
    def register_levels(self, blocks, block=1):
        """Register a block's multigrid levels, finest first."""
        for m, b in enumerate(blocks):
            self.register(b, level=m + 1, block=block)

    def restrict(self, mfine, block=1):
        self._chk(lib().ff_restrict(self.h, mfine, block), "ff_restrict")

    def forcing(self, level, phase, block=1):
        self._chk(lib().ff_forcing(self.h, level, block, phase), "ff_forcing")

    def prolong(self, mcoarse, block=1):
        self._chk(lib().ff_prolong(self.h, mcoarse, block), "ff_prolong")

    def vcycle(self, nlev, ncycles=1, block=1):
        resi = np.zeros(7)
        self._chk(lib().ff_vcycle(self.h, block, nlev, ncycles, resi.ctypes.data), "ff_vcycle")
        return resi

    def vcycle_graph(self, nlev, ncycles=1, block=1):
        """ncycles V-cycles replayed from one captured HIP graph (no residual sums)."""
        self._chk(lib().ff_vcycle_graph(self.h, block, nlev, ncycles), "ff_vcycle_graph")

    # ---- solid-wall patches (BOTSKE) ---------------------------------------------------------
    def wall_register(self, wparams, patches, level=1, block=1):
        """Register the solid-wall patches (finflo_amd.layout.WallPatch, reference order) of a
        block level; their host arrays are mirrored on the device and stay the caller's."""
        views = wall_views(patches)
        self._walls = getattr(self, "_walls", {})
        self._walls[(level, block)] = (wparams, patches, views)
        self._chk(lib().ff_wall_register(self.h, level, block, C.byref(wparams), views,
                                         len(patches)), "ff_wall_register")

    def wall_sync(self, direction, level=1, block=1):
        """direction 0: patch arrays host -> device, 1: device -> host."""
        self._chk(lib().ff_wall_sync(self.h, level, block, direction), "ff_wall_sync")

    def wall_forces(self, patch, block=1):
        """The 19 force sums (CX .. BUZ) of a patch from the last level-1 residual."""
        out = np.zeros(19)
        self._chk(lib().ff_wall_forces(self.h, block, patch, out.ctypes.data), "ff_wall_forces")
        return out

    # ---- the learned EARSM closure (device sample arrays are torch tensors) -------------
    def earsm_mode(self, mode):
        """EARSM1's coefficients: "exact" or "mlp" (the learned closure in its box)."""
        self.set_option(FF_OPT_EARSM, {"exact": FF_EARSM_EXACT, "mlp": FF_EARSM_MLP}[mode])

    def mlp_set(self, params=None, norm=None):
        p = None if params is None else np.ascontiguousarray(params, dtype=np.float32)
        q = None if norm is None else np.ascontiguousarray(norm, dtype=np.float32)
        self._chk(lib().ff_mlp_set(self.h, None if p is None else p.ctypes.data,
                                   None if q is None else q.ctypes.data), "ff_mlp_set")

    def mlp_get(self):
        p = np.zeros(FF_MLP_NPARAM, dtype=np.float32)
        q = np.zeros(FF_MLP_NNORM, dtype=np.float32)
        self._chk(lib().ff_mlp_get(self.h, p.ctypes.data, q.ctypes.data), "ff_mlp_get")
        return p, q

    @staticmethod
    def _dev(t, cols):
        import torch
        assert t.is_cuda and t.dtype == torch.float32 and t.is_contiguous() and t.shape[-1] == cols
        return t.data_ptr()

    def earsm_samples(self, level=1, block=1):
        """Exact-EARSM samples of every interior cell: (x (n, 5), y (n, 3)) on the device."""
        import torch
        blk = self._views[(level, block)][0]
        n = blk.imax * blk.jmax * blk.kmax
        x = torch.empty((n, 5), dtype=torch.float32, device="cuda")
        y = torch.empty((n, 3), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        self._chk(lib().ff_earsm_samples(self.h, level, block, x.data_ptr(), y.data_ptr()),
                  "ff_earsm_samples")
        return x, y

    def mlp_fit_norm(self, x, y):
        import torch
        torch.cuda.synchronize()
        self._chk(lib().ff_mlp_fit_norm(self.h, self._dev(x, 5), self._dev(y, 3), len(x)),
                  "ff_mlp_fit_norm")

    def mlp_infer(self, x):
        import torch
        y = torch.empty((len(x), 3), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        self._chk(lib().ff_mlp_infer(self.h, self._dev(x, 5), len(x), y.data_ptr()), "ff_mlp_infer")
        return y

    def mlp_grad(self, x, y):
        """(loss, gradient) over all samples."""
        import torch
        g = np.zeros(FF_MLP_NPARAM + 1, dtype=np.float32)
        torch.cuda.synchronize()
        self._chk(lib().ff_mlp_grad(self.h, self._dev(x, 5), self._dev(y, 3), len(x),
                                    g.ctypes.data), "ff_mlp_grad")
        return float(g[-1]), g[:-1]

    def mlp_train(self, x, y, steps, batch=4096, lr=1e-3, seed=0):
        """Adam steps on minibatches; returns the minibatch losses."""
        import torch
        loss = np.zeros(max(steps, 1), dtype=np.float32)
        torch.cuda.synchronize()
        self._chk(lib().ff_mlp_train(self.h, self._dev(x, 5), self._dev(y, 3), len(x), steps,
                                     batch, lr, seed, loss.ctypes.data), "ff_mlp_train")
        return loss[:steps]

    # ---- timing ------------------------------------------------------------------------
    def profile(self, on=True):
        self._chk(lib().ff_ctx_profile(self.h, 1 if on else 0), "ff_ctx_profile")

    def kernel_stats(self, name):
        n, ms = C.c_longlong(), C.c_double()
        self._chk(lib().ff_kernel_stats(self.h, name.encode(), C.byref(n), C.byref(ms)),
                  "ff_kernel_stats")
        return n.value, ms.value

    def reset_stats(self):
        self._chk(lib().ff_kernel_stats_reset(self.h), "ff_kernel_stats_reset")
