"""Every texture wrap and filter mode on the GPU against the oracle.

The reference's textures take ForkerGL's TextureWrapping / TextureFiltering
at load time (src/forkergl.cpp:15-16, src/model.cpp:425-426): NoWrap and
Nearest unless changed first -- the modes of every reference render.  Here
the scene carries no `texture` key and is loaded with each of the 4 x 2
modes; texture coordinates are scaled to run outside [0, 1] (negative and
above 1), so the wrap rule decides most samples.  G-buffer (albedo,
roughness, metalness and the TBN-mapped normal all come from textures),
depth and shadow map bit-exact; the 8-bit image within +-1 LSB.
"""
import os

import pytest

import forkerrenderer_amd as F
from forkerrenderer_amd import scenegen as G
from tests.util import assert_bitexact, gpu_frame, lsb_stats, oracle_frame

pytestmark = pytest.mark.gpu

WRAPS = {"none": F.WRAP_NONE, "repeat": F.WRAP_REPEAT, "mirror": F.WRAP_MIRROR, "clamp": F.WRAP_CLAMP}
FILTERS = {"nearest": F.FILTER_NEAREST, "linear": F.FILTER_LINEAR}


@pytest.mark.parametrize("mode", ["deferred", "forward"])
@pytest.mark.parametrize("filt", list(FILTERS))
@pytest.mark.parametrize("wrap", list(WRAPS))
def test_texture_modes(ctx, scene_dir, wrap, filt, mode):
    d = os.path.join(scene_dir, f"tex_{mode}")
    if not os.path.exists(os.path.join(d, "tiny.scene")):
        G.tiny_scene(d, seed=61, width=120, height=80, subdiv=3, tex=32, mode=mode, shadow="grid",
                     texture=None, uv_scale=3.0, pbr=(mode == "deferred"))
    sc = F.Scene(os.path.join(d, "tiny.scene"), WRAPS[wrap], FILTERS[filt])
    for i in range(sc.info.n_textures):
        t = sc.texture(i)
        assert (t["wrap"], t["filter"]) == (WRAPS[wrap], FILTERS[filt])
    print(f"texture modes: {wrap} {filt} {mode}", flush=True)
    got = gpu_frame(ctx, sc, 16)
    want = oracle_frame(sc)
    assert (want["depth"] < 3e38).sum() > 2000
    assert_bitexact("depth", got["depth"], want["depth"])
    assert_bitexact("shadow", got["shadow"], want["shadow"])
    if mode == "deferred":
        assert_bitexact("gbuf", got["gbuf"], want["gbuf"])
    mx, frac = lsb_stats(got["image"], want["image"])
    assert frac >= 0.999, (mx, frac)

