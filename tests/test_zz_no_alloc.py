"""The no-allocation merge (lnt_merge_into_noalloc) through the C-ABI against
lnt_merge_into, which it must equal bit for bit: two calls in a row on one
context, the second on smaller operands (the static scratch keeps the first
call's words), every operation, plus the size limit, rejected on the host."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pair(orc, na, nb, overlap, domain, seed):
    keys = orc.gen_distinct_keys(na + nb - overlap, domain, seed)
    ak = np.sort(keys[:na])
    bk = np.sort(np.concatenate([keys[:overlap], keys[na:]]))
    return ak, orc.gen_uniform_vals(na, seed + 1), bk, orc.gen_uniform_vals(len(bk), seed + 2)


def _merge(L, ctx, fn, ak, av, bk, bv, op, cap):
    import torch

    ck = torch.full((max(cap, 1),), -1, dtype=torch.int64, device="cuda")
    cv = torch.full((max(cap, 1),), -7.0, dtype=torch.float64, device="cuda")
    nc = ctypes.c_int64(-1)
    args = [ctx.handle, ak.data_ptr(), av.data_ptr(), ak.shape[0], bk.data_ptr(), bv.data_ptr(), bk.shape[0], op,
            ck.data_ptr(), cv.data_ptr(), cap, ctypes.byref(nc)]
    import libnt_amd

    if fn == "lnt_merge_into":
        args.append(libnt_amd.MEM_DEVICE)
    libnt_amd.check(getattr(L, fn)(*args))
    torch.cuda.synchronize()
    return ck[: nc.value].cpu().numpy(), cv[: nc.value].cpu().numpy()


def test_noalloc_merge_equals_merge_into(ctx, orc):
    import torch

    import libnt_amd

    L = libnt_amd.lib()
    # the second call is smaller: fewer tiles, scratch words left over from
    # the first call beyond them
    for call, (na, nb, ov, domain) in enumerate([(3_000_000, 2_500_000, 1_000_000, 1 << 40),
                                                 (70_000, 90_000, 30_000, 1 << 20)]):
        ak, av, bk, bv = _pair(orc, na, nb, ov, domain, 40 + call)
        g = [torch.from_numpy(x).cuda() for x in (ak, av, bk, bv)]
        for op in (0, 1, 2):
            print("no-alloc merge call", call, "sizes", na, nb, "op", op, flush=True)
            cap = min(na, nb) if op == 2 else na + nb
            ek, ev = _merge(L, ctx, "lnt_merge_into", *g, op, cap)
            k, v = _merge(L, ctx, "lnt_merge_into_noalloc", *g, op, cap)
            assert np.array_equal(k, ek) and np.array_equal(v.view(np.int64), ev.view(np.int64)), (call, op)
            assert len(k) == (ov if op == 2 else na + nb - ov)


def test_noalloc_merge_limits(ctx):
    import torch

    import libnt_amd

    L = libnt_amd.lib()
    k = torch.zeros(1, dtype=torch.int64, device="cuda")
    v = torch.zeros(1, dtype=torch.float64, device="cuda")
    nc = ctypes.c_int64()
    # Beyond the static scratch: na + nb above 43,000 tiles of 8,192 steps is
    # refused on the host before any launch (the arrays are never read).
    n = 43_000 * 8192 // 2 + 1
    rc = L.lnt_merge_into_noalloc(ctx.handle, k.data_ptr(), v.data_ptr(), n, k.data_ptr(), v.data_ptr(), n, 0,
                                  k.data_ptr(), v.data_ptr(), 2 * n, ctypes.byref(nc))
    assert rc == 6 and b"no-allocation" in L.lnt_last_error()
    # empty operands: nothing to do, nothing written
    rc = L.lnt_merge_into_noalloc(ctx.handle, None, None, 0, None, None, 0, 0, None, None, 0, ctypes.byref(nc))
    assert rc == 0 and nc.value == 0
