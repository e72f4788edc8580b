"""Edge cases of the frame path against the oracle: screen sizes that leave
partial tiles and rows that are not a whole number of 16-byte words (the
SSAO staging takes its per-float path), and a main pass whose triangles all
miss the screen (no (tile, triangle) pairs: the emit, sort and rasteriser
run on nothing and every pixel keeps its clear value)."""
import os

import pytest

import forkerrenderer_amd as F
from forkerrenderer_amd import scenegen as G
from tests.util import assert_bitexact, gpu_frame, lsb_stats, oracle_frame

pytestmark = pytest.mark.gpu


def _compare(ctx, p, tile):
    sc = F.Scene(p)
    print(f"scene {p} tile {tile}", flush=True)
    got = gpu_frame(ctx, sc, tile)
    want = oracle_frame(sc)
    assert_bitexact("depth", got["depth"], want["depth"])
    assert_bitexact("shadow", got["shadow"], want["shadow"])
    assert_bitexact("shadow_depth", got["shadow_depth"], want["shadow_depth"])
    assert_bitexact("gbuf", got["gbuf"], want["gbuf"])
    assert_bitexact("ao", got["ao"], want["ao"])
    mx, frac = lsb_stats(got["image"], want["image"])
    assert mx <= 1 or frac >= 0.999, (mx, frac)
    return want


@pytest.mark.parametrize("size", [(97, 61), (33, 17), (5, 3), (130, 7)], ids=lambda s: f"{s[0]}x{s[1]}")
@pytest.mark.parametrize("tile", [8, 32])
def test_odd_screen_sizes(ctx, scene_dir, size, tile):
    w, h = size
    p = G.tiny_scene(os.path.join(scene_dir, f"odd{w}x{h}"), seed=61, width=w, height=h, subdiv=2, tex=16,
                     ssao=True, shadow="pcss")
    want = _compare(ctx, p, tile)
    assert (want["depth"] < 3e38).any()


def test_main_pass_with_every_triangle_off_screen(ctx, scene_dir):
    # an orthographic camera centred 30 units to the side of the models
    p = G.tiny_scene(os.path.join(scene_dir, "offscreen"), seed=62, width=64, height=48, subdiv=2, tex=16,
                     ssao=True, shadow="grid", extra="camera ortho 30 0 5 30 0 0")
    want = _compare(ctx, p, 16)
    assert (want["depth"] >= 3e38).all(), "the scene is not off screen"
    assert (want["shadow_depth"] < 3e38).any(), "the shadow pass should still see the models"
