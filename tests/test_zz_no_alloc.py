"""fm_render_noalloc, the allocation-free megakernel render, against
fm_render on the same scene and random numbers: two calls in a row, the
second smaller, into caller-owned device arrays; and its size and mode
limits.  No counterpart in the reference (it serves callers that cannot
allocate device memory)."""
import ctypes as C

import numpy as np
import pytest

import fumeng_amd
from fumeng_amd import _abi as A
from fumeng_amd import scenes

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    """The HIP runtime libfumeng_gpu.so links (its NEEDED soname): torch
    ships another copy under the unversioned name, whose allocations the
    library's runtime would not know."""
    fumeng_amd.lib()
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return h


def test_noalloc_render_equals_fm_render(gpu_ctx):
    hip = _hip()
    s = gpu_ctx.scene(scenes.box_scene(width=64, height=64, spp=8))
    L = fumeng_amd.lib()
    nbytes = 64 * 64 * 3 * 4
    bufs = []
    try:
        for _ in range(3):
            p = C.c_void_p()
            assert hip.hipMalloc(C.byref(p), nbytes) == 0
            bufs.append(p)
        # the second call is smaller, and its 4-sample chunks add two partial sums per pixel
        for (W, H, spp, chunk, seed) in ((64, 64, 8, 8, 11), (40, 24, 8, 4, 12)):
            print("noalloc case", W, H, spp, chunk, seed, flush=True)
            ref = s.render(width=W, height=H, spp=spp, seed=seed, mode=A.FM_MODE_MEGAKERNEL, chunk=chunk)
            prm = A.fm_render_params(width=W, height=H, spp=spp, seed=seed, mode=A.FM_MODE_MEGAKERNEL, chunk=chunk,
                                     rr_depth=-1)
            st = A.fm_render_stats()
            rc = L.fm_render_noalloc(gpu_ctx.h, s.h, C.byref(prm), None, bufs[0], bufs[1], bufs[2], C.byref(st))
            assert rc == A.FM_OK, L.fm_last_error(gpu_ctx.h)
            assert st.paths == W * H * spp
            assert st.rays == ref[4].rays
            for k in range(3):
                got = np.empty((H, W, 3), np.float32)
                assert hip.hipMemcpy(got.ctypes.data, bufs[k], got.nbytes, D2H) == 0
                np.testing.assert_array_equal(got, ref[k])
    finally:
        for p in bufs:
            hip.hipFree(p)
        s.close()


def test_noalloc_render_rejects_what_it_cannot_hold(gpu_ctx):
    s = gpu_ctx.scene(scenes.box_scene(width=8, height=8, spp=1))
    L = fumeng_amd.lib()
    try:
        side = int(np.sqrt(A.FM_NOALLOC_MAX_PIXELS)) + 1
        for prm in (A.fm_render_params(width=side, height=side, spp=1, mode=A.FM_MODE_MEGAKERNEL, rr_depth=-1),
                    A.fm_render_params(width=8, height=8, spp=1, mode=A.FM_MODE_WAVEFRONT, rr_depth=-1),
                    A.fm_render_params(width=8, height=8, spp=1, mode=A.FM_MODE_MEGAKERNEL, accumulate=1,
                                       rr_depth=-1)):
            rc = L.fm_render_noalloc(gpu_ctx.h, s.h, C.byref(prm), None, None, None, None, None)
            assert rc == A.FM_ERR_INVALID
            assert L.fm_last_error(gpu_ctx.h)
    finally:
        s.close()
