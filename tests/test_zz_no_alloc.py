"""The allocation-free entry point vxhip_std_sort_noalloc (caller-owned device
buffers, a fixed per-GPU working array) against the allocating entry point
vxhip_probe_std_sort on the same inputs: identical order, on two calls in a
row (the second smaller, with segments), then a size past the limit and
segment offsets that fall, both refused.  Needs an MI355X."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(ctx, keys, vals, seg=None):
    """Sorts copies of keys/vals in device memory through the C ABI."""
    import torch
    from velox_amd.device import library
    k = torch.from_numpy(keys.view(np.int64).copy()).to("cuda")
    v = torch.from_numpy(vals.view(np.int32).copy()).to("cuda")
    s = None if seg is None else torch.from_numpy(np.asarray(seg, np.uint32).view(np.int32).copy()).to("cuda")
    torch.cuda.synchronize()
    rc = library().vxhip_std_sort_noalloc(ctx._h, C.c_void_p(k.data_ptr()), C.c_void_p(v.data_ptr()), len(keys),
                                          C.c_void_p(s.data_ptr()) if s is not None else None,
                                          0 if seg is None else len(seg) - 1)
    return rc, k.cpu().numpy().view(np.uint64), v.cpu().numpy().view(np.uint32)


def test_noalloc_sort_matches_allocating_entry_point():
    import torch  # noqa: F401  (loaded before the library, so both share torch's HIP runtime)
    from velox_amd.device import Context, ERR_CAPACITY, ERR_INVALID, OK
    ctx = Context()
    try:
        rng = np.random.default_rng(11)
        # First call: one segment of 200,000 keys with many ties.
        keys = rng.integers(0, 5000, 200000).astype(np.uint64)
        vals = np.arange(len(keys), dtype=np.uint32)
        print("[gpu] no-alloc sort: 200000 keys, 1 segment", flush=True)
        rc, gk, gv = _noalloc(ctx, keys, vals)
        assert rc == OK
        wk, wv = ctx.std_sort(keys, vals)
        assert np.array_equal(gk, wk) and np.array_equal(gv, wv)
        # Second call, smaller, with 37 segments (the array keeps the first
        # call's contents: everything the sort needs zeroed is cleared).
        keys2 = rng.integers(0, 300, 50000).astype(np.uint64)
        vals2 = rng.permutation(50000).astype(np.uint32)
        seg = np.concatenate([[0], np.sort(rng.integers(0, 50000, 36)), [50000]]).astype(np.uint32)
        print("[gpu] no-alloc sort: 50000 keys, 37 segments", flush=True)
        rc, gk, gv = _noalloc(ctx, keys2, vals2, seg)
        assert rc == OK
        wk, wv = ctx.std_sort(keys2, vals2, seg)
        assert np.array_equal(gk, wk) and np.array_equal(gv, wv)
        # Past the limit: refused before any device work.
        big = np.zeros(262145, np.uint64)
        rc, _, _ = _noalloc(ctx, big, np.zeros(len(big), np.uint32))
        assert rc == ERR_CAPACITY
        # Offsets that fall: refused, keys left as they were.
        bad = np.array([0, 30000, 20000, 50000], np.uint32)
        rc, gk, gv = _noalloc(ctx, keys2, vals2, bad)
        assert rc == ERR_INVALID
        assert np.array_equal(gk, keys2) and np.array_equal(gv, vals2)
    finally:
        ctx.close()
