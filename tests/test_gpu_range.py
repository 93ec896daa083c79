"""The r2r value range computed on the device (k_range partials reduced by
k_ebfam) equals the oracle's GetDatumValueRange restatement
(src/types.cc:27-35): on a field large enough that every workgroup slot
covers many strides with the extremes far apart, through the scalar path of
a field that is not 16-byte aligned, and with interior NaNs (ignored, as
std::min_element / std::max_element ignore them after the first element).
The range is read back from the context's metadata record."""
import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _device_range(ctx, xt):
    from vecsz_amd import _lib

    print("range", tuple(xt.shape), xt.data_ptr() % 16, flush=True)
    ctx.compress(xt, 1e-3, "r2r")
    return _lib.Sidecar.from_buffer_copy(ctx.last_sidecar()).value_range


def test_device_range_matches_oracle(gpu_ctx):
    rng = np.random.default_rng(21)
    n = (1 << 24) + 5
    x = rng.normal(0, 1, n).astype(np.float32)
    x[3] = 97.5
    x[n - 2] = -88.25
    x[n // 2 + 1] = 50.0
    xt = torch.from_numpy(x).cuda()
    assert _device_range(gpu_ctx, xt) == po.value_range(x)

    # not 16-byte aligned: element loads
    buf = torch.zeros(n + 1, dtype=torch.float32, device="cuda")
    buf[1:] = xt
    assert _device_range(gpu_ctx, buf[1:]) == po.value_range(x)

    # the maximum in the field's last, partial float4 group
    y = x.copy()
    y[n - 1] = 123.0
    assert _device_range(gpu_ctx, torch.from_numpy(y).cuda()) == po.value_range(y)

    # interior NaNs are skipped
    z = rng.normal(0, 1, 100_003).astype(np.float32)
    z[100] = np.nan
    z[77_777] = np.nan
    z[5] = -9.0
    got = _device_range(gpu_ctx, torch.from_numpy(z).cuda())
    assert got == po.value_range(z) and got == float(np.float32(np.nanmax(z)) - np.float32(np.nanmin(z)))
