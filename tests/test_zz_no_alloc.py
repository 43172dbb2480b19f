"""binder_fold_mask_3d_noalloc (device inputs and outputs, scratch in a
file-scope device array, nothing allocated) against binder_fold_mask_3d,
through the C ABI: the same fold mask and count, bit for bit, on two calls
in a row, the second with a smaller input."""
import ctypes
import os

import numpy as np
import pytest

_c_int, _c_double, _vp = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
HOST_TO_DEVICE, DEVICE_TO_HOST = 1, 2


def _hip():
    """The HIP runtime the library itself links (test-side buffers only)."""
    hip = ctypes.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib",
                                   "libamdhip64.so"))
    hip.hipMalloc.argtypes = [ctypes.POINTER(_vp), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [_vp, _vp, ctypes.c_size_t, _c_int]
    hip.hipFree.argtypes = [_vp]
    hip.hipStreamCreate.argtypes = [ctypes.POINTER(_vp)]
    hip.hipStreamSynchronize.argtypes = [_vp]
    hip.hipStreamDestroy.argtypes = [_vp]
    return hip


def _folded_field(shape, seed, amp):
    """Positions plus a smooth random displacement strong enough to fold."""
    rng = np.random.default_rng(seed)
    pos = np.stack(np.indices(shape, dtype=np.float64), -1)
    disp = np.zeros_like(pos)
    for _ in range(6):
        k = rng.uniform(0.2, 0.6, 3)
        ph = rng.uniform(0, 2 * np.pi, 3)
        a = rng.normal(0, amp, 3)
        wave = np.cos(pos[..., 0] * k[0] + ph[0]) * np.cos(pos[..., 1] * k[1] + ph[1]) * \
            np.cos(pos[..., 2] * k[2] + ph[2])
        disp += wave[..., None] * a
    return pos + disp


@pytest.mark.gpu
def test_fold_mask_noalloc_matches_allocating_entry(core):
    print("fold mask without allocation vs binder_fold_mask_3d", flush=True)
    hip = _hip()
    fn = core._lib.binder_fold_mask_3d_noalloc
    fn.restype = _c_int
    fn.argtypes = [_vp, _c_int, _c_int, _c_int, _c_double, _c_double, _c_double, _c_double,
                   _vp, _vp, _vp]
    vs = (1.0, 1.2, 0.9)
    stream = _vp()
    assert hip.hipStreamCreate(ctypes.byref(stream)) == 0
    bufs = []

    def dmalloc(n):
        p = _vp()
        assert hip.hipMalloc(ctypes.byref(p), n) == 0
        bufs.append(p)
        return p

    try:
        # the second call is smaller, on the null stream; the scratch keeps
        # the first call's partials, which must not leak into its result
        for shape, seed, st in (((48, 40, 36), 1, stream), ((19, 13, 11), 2, None)):
            f = _folded_field(shape, seed, 2.5)
            nv = int(np.prod(shape))
            mask_ref, cnt_ref = core.fold_mask_3D(f.ravel(), *shape, *vs, e1=0.01)
            assert 0 < cnt_ref < nv
            planes = np.ascontiguousarray(np.moveaxis(f, -1, 0))
            d_f, d_m, d_c = dmalloc(planes.nbytes), dmalloc(nv), dmalloc(8)
            assert hip.hipMemcpy(d_f, planes.ctypes.data, planes.nbytes, HOST_TO_DEVICE) == 0
            junk = np.full(1, -7, np.int64)
            assert hip.hipMemcpy(d_c, junk.ctypes.data, 8, HOST_TO_DEVICE) == 0
            rc = fn(d_f, *shape, *vs, 0.01, d_m, d_c, st)
            assert rc == 0, core._lib.binder_last_error()
            assert hip.hipStreamSynchronize(st) == 0
            mask = np.empty(nv, np.uint8)
            cnt = np.empty(1, np.int64)
            assert hip.hipMemcpy(mask.ctypes.data, d_m, nv, DEVICE_TO_HOST) == 0
            assert hip.hipMemcpy(cnt.ctypes.data, d_c, 8, DEVICE_TO_HOST) == 0
            assert int(cnt[0]) == cnt_ref, (shape, int(cnt[0]), cnt_ref)
            assert np.array_equal(mask.reshape(shape).astype(bool), mask_ref), shape
            # a NULL mask still gives the count
            rc = fn(d_f, *shape, *vs, 0.01, None, d_c, st)
            assert rc == 0
            assert hip.hipStreamSynchronize(st) == 0
            assert hip.hipMemcpy(cnt.ctypes.data, d_c, 8, DEVICE_TO_HOST) == 0
            assert int(cnt[0]) == cnt_ref
        # more than 2^31 - 1 voxels is refused before any launch
        rc = fn(bufs[0], 2048, 1024, 1024, *vs, 0.01, None, bufs[2], None)
        assert rc == 1 and b"2^31" in core._lib.binder_last_error()
    finally:
        for p in bufs:
            hip.hipFree(p)
        hip.hipStreamDestroy(stream)
