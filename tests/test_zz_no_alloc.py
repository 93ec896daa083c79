"""vecsz_compress_noalloc through the C ABI: the archive is vecsz_compress's
byte for byte (and the oracle's), on two calls in a row with the second
smaller, so that what the first call left in the shared scratch array
cannot leak into the second; sizes beyond the stated limits are refused."""
import ctypes

import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _call(ctx, fn, p, x, cap):
    import torch

    out = torch.empty(cap, dtype=torch.uint8, device="cuda")
    size, nout = ctypes.c_uint64(), ctypes.c_uint64()
    torch.cuda.synchronize()
    rc = fn(ctx._ctx, ctypes.byref(p), ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(out.data_ptr()), cap,
            ctypes.byref(size), ctypes.byref(nout))
    return rc, out[:size.value].cpu().numpy().tobytes(), nout.value


@pytest.mark.parametrize("shapes,eb,mode", [
    (((160, 128, 192), (40, 30, 20)), 1e-4, "r2r"),      # 3D, 3.9 Mi then 24 Ki elements
    (((2048, 2048), (300, 77)), 1e-3, "abs"),           # 2D at the 2^22 limit, then a ragged field
    (((1 << 22,), (100003,)), 1e-4, "r2r"),             # 1D at the limit, then a partial last block
])
def test_noalloc_matches_compress_twice(gpu_ctx, shapes, eb, mode):
    import torch

    from vecsz_amd import datagen, make_params
    from vecsz_amd import _lib

    lib = _lib.load()
    for shape in shapes:
        print(f"noalloc: shape {shape} eb {eb} {mode}", flush=True)
        if len(shape) == 1:
            x = datagen.walk_spikes_1d(shape[0], seed=3)
        elif len(shape) == 2:
            x = datagen.grf(shape, seed=3)
        else:
            x = datagen.grf(shape, seed=3)
        x = np.ascontiguousarray(x, dtype=np.float32)
        xd = torch.from_numpy(x).cuda()
        p = make_params(tuple(reversed(x.shape)), eb, mode)
        cap = gpu_ctx.compress_bound(p)
        rc, a_na, n_na = _call(gpu_ctx, lib.vecsz_compress_noalloc, p, xd, cap)
        assert rc == 0, gpu_ctx._lib.vecsz_ctx_last_error(gpu_ctx._ctx)
        rc, a, n = _call(gpu_ctx, lib.vecsz_compress, p, xd, cap)
        assert rc == 0
        assert a_na == a and n_na == n
        want, n_ref, _ = po.compress(x, eb, mode == "r2r")
        assert a_na == want and n_na == n_ref


def test_noalloc_refuses_beyond_its_limits(gpu_ctx):
    import torch

    from vecsz_amd import make_params
    from vecsz_amd import _lib

    lib = _lib.load()
    x = torch.zeros((1 << 22) + 1, dtype=torch.float32, device="cuda")
    for p in (make_params(((1 << 22) + 1,), 1e-4, "abs"),                # one element too many
              make_params((1024,), 1e-4, "abs", dict_size=8192),         # dictionary too large
              make_params((1024,), 1e-4, "abs", pad_mode="block")):      # padding
        rc, _, _ = _call(gpu_ctx, lib.vecsz_compress_noalloc, p, x, 1 << 20)
        assert rc == 1, rc  # VECSZ_EINVAL
    # the context still works afterwards
    p = make_params((1024,), 1e-4, "abs")
    rc, a, _ = _call(gpu_ctx, lib.vecsz_compress_noalloc, p, x, 1 << 20)
    assert rc == 0 and len(a) > 0
