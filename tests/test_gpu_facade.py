"""The C++ facade (include/forkergl_amd.hpp) driving a frame the way the
reference's Render::Render does (ForkerGL statics, DepthShader / GShader /
BlinnPhongShader uniforms, Model::Render; forkerrenderer_amd/csrc/
facade_render.cpp) gives the same targets, bit for bit, as fgl_render_scene,
whose output the parity tests check against the oracle."""
import os
import subprocess

import numpy as np
import pytest

import forkerrenderer_amd as F
from forkerrenderer_amd import scenegen as G
from tests.util import assert_bitexact, gpu_frame

EXE = os.path.join(os.path.dirname(F.__file__), "_lib", "facade_render")


@pytest.mark.gpu
@pytest.mark.parametrize("kw", [dict(mode="deferred", ssao=True, shadow="pcss", ssaa=2),
                                dict(mode="forward", shadow="pcf"),
                                dict(mode="deferred", ssao=True, shadow="hard", pbr=False)],
                         ids=["deferred_pcss_ssaa", "forward_pcf", "deferred_phong"])
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
    rd = lambda n, t: np.fromfile(os.path.join(out, n), dtype=t)
    assert_bitexact("depth", rd("depth.f32", np.float32), want["depth"])
    assert_bitexact("fb", rd("fb.f32", np.float32), want["fb"].reshape(-1))
    assert np.array_equal(rd("image.u8", np.uint8), want["image"].reshape(-1))
    if sc.info.deferred:
        assert_bitexact("gbuf", rd("gbuf.f32", np.float32), want["gbuf"].reshape(-1))
        assert_bitexact("ao", rd("ao.f32", np.float32), want["ao"])


def test_facade_program_usage():
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "Usage" in r.stderr
