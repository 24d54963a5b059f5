"""ap_relmse_f64_noalloc (partials in the library's static device array)
against ap_relmse_f64 (partials in a buffer the library allocates), called
through the C ABI: bit-identical results on two calls in a row, the second
on a smaller input, so a partial left over from the first call would show;
and sizes beyond the limit rejected before any launch."""
import ctypes

import numpy as np
import pytest

import aperture as ap

pytestmark = pytest.mark.gpu


def _call(fn, img, ref, n, eps, out, stream):
    ap.check(fn(ctypes.byref(img), ctypes.byref(ref), n, eps, ap._ptr(out), ap._stream(stream)))


def test_noalloc_relmse_equals_the_allocating_entry_point(gpu):
    L = ap.lib()
    g = np.random.default_rng(12)
    stream = ap.Stream()
    for n, kind in ((700_001, ap.IMAGE_XYZ_F32), (3_001, ap.IMAGE_XYZ_SUMS_F64)):
        print(f"[no_alloc] n {n}, image kind {kind}", flush=True)
        ref = g.gamma(1.0, 0.5, (n, 3)).astype(np.float32)
        img = ref * g.uniform(0.5, 1.5, ref.shape)
        img = img.astype(np.float32) if kind == ap.IMAGE_XYZ_F32 else img * 8.0
        d_ref, d_img = ap.DeviceBuffer.from_numpy(ref), ap.DeviceBuffer.from_numpy(img)
        iv = ap.ImageView.of(d_img, kind, 1.0 if kind == ap.IMAGE_XYZ_F32 else 0.125)
        rv = ap.ImageView.of(d_ref)
        out_a, out_b = ap.DeviceBuffer(8), ap.DeviceBuffer(8)
        _call(L.ap_relmse_f64, iv, rv, n, 1e-2, out_a, stream)
        _call(L.ap_relmse_f64_noalloc, iv, rv, n, 1e-2, out_b, stream)
        stream.synchronize()
        a = out_a.to_numpy(np.float64, 1)[0]
        b = out_b.to_numpy(np.float64, 1)[0]
        print(f"[no_alloc] allocating {a:.17g}, no-alloc {b:.17g}", flush=True)
        assert np.isfinite(a) and a > 0
        assert np.float64(a).view(np.uint64) == np.float64(b).view(np.uint64)
    big = ap.DeviceBuffer(16)
    with pytest.raises(ap.ApertureError) as e:
        _call(L.ap_relmse_f64_noalloc, ap.ImageView.of(big), ap.ImageView.of(big), ap.REDUCE_MAX_ELEMENTS + 1, 1e-2,
              big, stream)
    assert e.value.code == ap.ERR_INVALID_ARG
