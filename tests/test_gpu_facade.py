"""The C++ facade (include/forkergl_amd.hpp) as a drop-in for the reference's
ForkerGL / shader / Scene interfaces.

* oracle/_ref/render_port is the reference's own frame-driver code
  (src/render.cpp: Preconfigure, Render, DoShadowPass, DoForwardPass,
  DoGeometryPass, DoLightingPass) compiled unchanged against the facade
  (oracle/ref_render_port.py; built where the reference tree is present and
  carried with the tree).  Its frames equal fgl_render_scene's bit for bit
  -- and fgl_render_scene's frames are checked against the oracle by the
  parity tests.
* forkerrenderer_amd/_lib/facade_render is this repository's own frame
  driver written against the facade, with the shadow-map-size extension and
  ForkerGL::TextureWrapMode / TextureFilterMode set before the scene loads.
"""
import os
import subprocess

import numpy as np
import pytest

import forkerrenderer_amd as F
from forkerrenderer_amd import scenegen as G
from tests.util import assert_bitexact, gpu_frame

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(os.path.dirname(F.__file__), "_lib", "facade_render")
PORT = os.path.join(ROOT, "oracle", "_ref", "render_port")


def _compare(out, want, deferred, shadow):
    rd = lambda n, t: np.fromfile(os.path.join(out, n), dtype=t)
    assert_bitexact("depth", rd("depth.f32", np.float32), want["depth"])
    assert_bitexact("fb", rd("fb.f32", np.float32), want["fb"].reshape(-1))
    assert np.array_equal(rd("image.u8", np.uint8), want["image"].reshape(-1))
    if shadow:
        assert_bitexact("shadow", rd("shadow.f32", np.float32), want["shadow"])
    if deferred:
        assert_bitexact("gbuf", rd("gbuf.f32", np.float32), want["gbuf"].reshape(-1))
        assert_bitexact("ao", rd("ao.f32", np.float32), want["ao"])


CASES = [dict(mode="deferred", ssao=True, shadow="pcss", ssaa=2),
         dict(mode="forward", shadow="pcf"),
         dict(mode="forward", shadow="hard", pbr=False),
         dict(mode="deferred", ssao=True, shadow="hard", pbr=False)]
IDS = ["deferred_pcss_ssaa", "forward_pcf", "forward_phong", "deferred_phong"]


@pytest.mark.gpu
@pytest.mark.parametrize("kw", CASES, ids=IDS)
def test_reference_pass_code_equals_render_scene(ctx, scene_dir, kw):
    """The reference's Render::* compiled against the facade.  Scenes keep the
    reference's frame-sized shadow map and load-time texture defaults."""
    if not os.path.exists(PORT):
        pytest.skip("oracle/_ref/render_port not built (the reference tree was absent at build time)")
    d = os.path.join(scene_dir, "port_" + "_".join(str(v) for v in kw.values()))
    p = G.tiny_scene(d, seed=52, width=80, height=56, subdiv=3, tex=16, texture=None, shadowmap=None, **kw)
    out = os.path.join(d, "out")
    os.makedirs(out, exist_ok=True)
    print(f"reference pass code through the facade: {kw}", flush=True)
    r = subprocess.run([PORT, p, out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    sc = F.Scene(p)
    _compare(out, gpu_frame(ctx, sc), sc.info.deferred, sc.info.shadow_on)


@pytest.mark.gpu
@pytest.mark.parametrize("kw", CASES, ids=IDS)
def test_facade_frame_equals_render_scene(ctx, scene_dir, kw):
    d = os.path.join(scene_dir, "facade_" + "_".join(str(v) for v in kw.values()))
    p = G.tiny_scene(d, seed=51, width=72, height=52, subdiv=3, tex=16, **kw)
    out = os.path.join(d, "out")
    os.makedirs(out, exist_ok=True)
    print(f"facade frame {kw}", flush=True)
    r = subprocess.run([EXE, p, out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    sc = F.Scene(p)
    want = gpu_frame(ctx, sc)
    _compare(out, want, sc.info.deferred, sc.info.shadow_on)
    if sc.info.deferred:   # a named G-buffer is its planes of the G-buffer target
        albedo = np.fromfile(os.path.join(out, "albedo.f32"), np.float32)
        assert_bitexact("AlbedoGBuffer", albedo, want["gbuf"][9:12].reshape(-1))


@pytest.mark.gpu
@pytest.mark.parametrize("wrap,filt", [(F.WRAP_MIRROR, F.FILTER_LINEAR), (F.WRAP_CLAMP, F.FILTER_NEAREST)])
def test_facade_texture_modes_apply_at_load(ctx, scene_dir, wrap, filt):
    """ForkerGL::TextureWrapMode / TextureFilterMode set before the scene is
    loaded give the frame fgl_render_scene renders from the scene loaded
    with those modes."""
    d = os.path.join(scene_dir, f"facade_tex_{wrap}_{filt}")
    p = G.tiny_scene(d, seed=53, width=72, height=52, subdiv=3, tex=16, texture=None, uv_scale=3.0,
                     mode="deferred", ssao=True)
    out = os.path.join(d, "out")
    os.makedirs(out, exist_ok=True)
    r = subprocess.run([EXE, p, out, str(wrap), str(filt)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    sc = F.Scene(p, wrap, filt)
    _compare(out, gpu_frame(ctx, sc), True, True)


def test_facade_program_usage():
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "Usage" in r.stderr
