"""fgl_ssao_noalloc -- the allocation-free SSAO entry point -- against
fgl_ssao on the same frames, through the C ABI.

The frame is rendered with fgl_render_scene (whose fgl_ssao writes the AO
target); its depth and G-buffer are copied into device buffers of the
test's own and fgl_ssao_noalloc runs on them.  Two frames in a row, the
second smaller, for each blur mode (none, the reference's recursive blur,
the separable sigma=2 blur whose unblurred AO lives in the static scratch
array).  AO bit-exact.
"""
import ctypes as C
import os

import numpy as np
import pytest

import forkerrenderer_amd as F
from forkerrenderer_amd import scenegen as G

pytestmark = pytest.mark.gpu

MAX_PIXELS = 1 << 23


def _hip():
    """The HIP runtime the library itself is linked against (device buffers
    for the test's inputs and outputs)."""
    F.lib()
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


H2D, D2H = 1, 2


class DevBuf:
    def __init__(self, hip, nbytes):
        self.hip, self.p = hip, C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), nbytes) == 0
        self.nbytes = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, n, dtype=np.float32):
        out = np.empty(n, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        if self.p:
            self.hip.hipFree(self.p)
            self.p = C.c_void_p()


def _bind():
    L = F.lib()
    f = L.fgl_ssao_noalloc
    P = C.c_void_p
    f.restype = C.c_int
    f.argtypes = [P, C.c_int, C.c_int, P, P, P, P, P, C.POINTER(F.SsaoCfg)]
    return f


def _camera(I):
    ratio = np.float32(I.width) / np.float32(I.height)
    V = F.make_lookat(list(I.eye), list(I.look_at))
    P = F.make_perspective(45.0, float(ratio), 0.01, 20.0)
    return np.ascontiguousarray(V, np.float32).reshape(16), np.ascontiguousarray(P, np.float32).reshape(16)


@pytest.mark.parametrize("blur", ["gaussian 2", "reference", "none"])
def test_noalloc_ssao_equals_fgl_ssao(ctx, scene_dir, blur):
    hip = _hip()
    fn = _bind()
    # sized for the larger frame; the smaller one reuses the same buffers
    cap = 200 * 136
    d_depth, d_gbuf, d_ao = DevBuf(hip, cap * 4), DevBuf(hip, F.GBUF_PLANES * cap * 4), DevBuf(hip, cap * 4)
    try:
        for i, (w, h) in enumerate(((200, 136), (88, 60))):
            p = G.tiny_scene(os.path.join(scene_dir, f"noalloc_{blur.split()[0]}_{i}"), seed=40 + i, width=w,
                             height=h, subdiv=3, tex=16, ssao=True, extra=f"aoblur {blur}")
            sc = F.Scene(p)
            I = sc.info
            print(f"noalloc: blur {blur}, frame {w}x{h}", flush=True)
            ctx.render_scene(sc)
            depth = ctx.read(F.TARGET_DEPTH)
            gbuf = ctx.read(F.TARGET_GBUFFER)
            want = ctx.read(F.TARGET_AO)
            assert (want < 1).mean() > 0.05
            d_depth.put(depth)
            d_gbuf.put(gbuf)
            d_ao.put(np.full(w * h, np.nan, np.float32))
            V, Pm = _camera(I)
            rc = fn(ctx._h, w, h, V.ctypes.data, Pm.ctypes.data, d_depth.p, d_gbuf.p, d_ao.p, C.byref(I.ssao))
            assert rc == 0, F.lib().fgl_last_error(ctx._h)
            ctx.sync()
            got = d_ao.get(w * h)
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), \
                f"frame {w}x{h}: {(got != want).sum()} of {w * h} AO values differ"
    finally:
        for b in (d_depth, d_gbuf, d_ao):
            b.free()


def test_noalloc_ssao_rejects_what_it_cannot_run(ctx):
    fn = _bind()
    cfg = F.SsaoCfg()
    cfg.n_kernel, cfg.radius, cfg.blur = 16, 0.5, 0
    eye = np.eye(4, dtype=np.float32).reshape(16)
    dummy = C.c_void_p(16)   # never dereferenced: every case fails its host-side checks first
    assert fn(ctx._h, 4096, 2049, eye.ctypes.data, eye.ctypes.data, dummy, dummy, dummy, C.byref(cfg)) == \
        F.ERR_CAPACITY
    assert b"FGL_SSAO_NOALLOC_MAX_PIXELS" in F.lib().fgl_last_error(ctx._h)
    cfg.stochastic = 1
    assert fn(ctx._h, 64, 64, eye.ctypes.data, eye.ctypes.data, dummy, dummy, dummy, C.byref(cfg)) == \
        F.ERR_INVALID
    cfg.stochastic, cfg.n_kernel = 0, 65
    assert fn(ctx._h, 64, 64, eye.ctypes.data, eye.ctypes.data, dummy, dummy, dummy, C.byref(cfg)) == \
        F.ERR_INVALID
    assert fn(ctx._h, 64, 64, eye.ctypes.data, eye.ctypes.data, None, dummy, dummy, C.byref(cfg)) == F.ERR_INVALID
