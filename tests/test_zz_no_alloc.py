"""ff_residual_noalloc (every scratch buffer in the library's fixed device array, no
registration, no allocation) against ff_residual on a registered block, through the C-ABI:
two calls in a row on the same context, the second on a smaller block, bit for bit in
every field and in the RESL2 sums; and the size limit of the fixed array."""
import ctypes as C

import numpy as np
import pytest

from finflo_amd.cases import make_block
from finflo_amd.layout import NF, default_params

pytestmark = pytest.mark.gpu


def _device_copy(blk):
    import torch
    return torch.from_numpy(blk.F).cuda().contiguous()


@pytest.mark.parametrize("iturb", [0, 6], ids=["laminar", "sst"])
def test_residual_noalloc_matches_registered(hip_device, iturb):
    import torch
    from finflo_amd.ffhip import Context, lib
    prm = default_params(iturb=iturb)
    ctx = Context(prm)
    L = lib()
    for call, (dims, seed) in enumerate((((24, 16, 12), 3), ((12, 8, 6), 4))):
        print(f"case no-alloc residual iturb={iturb} call {call} dims={dims}", flush=True)
        blk = make_block(*dims, prm, distort=0.1, seed=seed, amp=0.03, kind="all",
                         turb=(1.0, 150.0) if iturb else None)
        t1 = _device_copy(blk)
        torch.cuda.synchronize()
        v1 = blk.view_struct(base=t1.data_ptr())
        assert L.ff_block_register_device(ctx.h, 1, 7, C.byref(v1)) == 0, ctx
        assert L.ff_state(ctx.h, 1, 7) == 0
        assert L.ff_turb(ctx.h, 1, 7) == 0
        assert L.ff_finish(ctx.h) == 0
        t2 = t1.clone()
        torch.cuda.synchronize()
        r1, r2 = np.zeros(7), np.zeros(7)
        assert L.ff_residual(ctx.h, 1, 7, 1, r1.ctypes.data) == 0
        v2 = blk.view_struct(base=t2.data_ptr())
        rc = L.ff_residual_noalloc(ctx.h, 1, C.byref(v2), 1, r2.ctypes.data)
        assert rc == 0, L.ff_last_error(ctx.h)
        assert L.ff_finish(ctx.h) == 0
        a, b = t1.cpu().numpy(), t2.cpu().numpy()
        bad = [i for i in range(NF) if not np.array_equal(a[i], b[i])]
        assert not bad, f"fields differ: {bad}"
        np.testing.assert_array_equal(r2, r1)
        assert np.all(r1[:5] > 0)
    # beyond the fixed array: a k-omega block of more than 8,388,608 entries is refused
    # before anything runs (the pointers are never read)
    if iturb:
        big = blk.view_struct(base=t2.data_ptr())
        big.imax, big.jmax, big.kmax = 300, 200, 150
        r = np.zeros(7)
        assert L.ff_residual_noalloc(ctx.h, 1, C.byref(big), 1, r.ctypes.data) == -1
        assert b"8388608" in L.ff_last_error(ctx.h)
    ctx.close()
