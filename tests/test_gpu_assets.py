"""The reference's own sample assets (obj/diablo_pose, mary, great_sword,
african_head, brickwall, plane; kept byte for byte under tests/golden/assets)
through the whole device path: the product's loader, shadow pass, forward
Blinn-Phong or deferred G-buffer + SSAO + blur + lighting, SSAA resolve.

Without filter keys the scenes take the reference's own filters: PCSS and
DoSSAO drawing from the random stream.  These exercise what the generated
scenes do not: mary's two groups drawn in
name order with a material whose texture file is missing, great_sword's
quads (fan triangulation) and RGBA texture, the head and its eye model with
1024^2 diffuse / specular / normal maps, brickwall's normal map with
generated tangents, Diablo's 5,022 faces (the README's timing scene).

Oracle: the same frame rendered by the oracle from the same loaded arrays;
coverage, depth, shadow map, G-buffer and AO bit-exact, the 8-bit image within
1 LSB on at least 99.9% of pixels.
"""
import os

import numpy as np
import pytest

import forkerrenderer_amd as F
import oracle as O
from tests import assets as A
from tests.util import assert_bitexact, gpu_frame, lsb_stats, oracle_frame

pytestmark = pytest.mark.gpu

SETS = [["plane", "diablo_pose"], ["plane", "mary"], ["plane", "great_sword"],
        ["african_head", "african_head/eye"], ["brickwall"]]


def _check(ctx, p, label, min_cover):
    sc = F.Scene(p)
    I = sc.info
    print(f"{label}: render on the device", flush=True)
    got = gpu_frame(ctx, sc)
    want = oracle_frame(sc)
    cover = float((want["depth"] < 3e38).mean())
    assert cover > min_cover, cover
    keys = ["depth"] + (["shadow", "shadow_depth"] if I.shadow_on else []) + (["gbuf", "ao"] if I.deferred else [])
    for k in keys:
        assert_bitexact(f"{label} {k}", got[k], want[k])
    mx, frac = lsb_stats(got["image"], want["image"])
    print(f"{label}: coverage {cover:.3f}, image max diff {mx}, within 1 LSB {frac:.6f}", flush=True)
    assert frac >= 0.999, (mx, frac)
    return sc


@pytest.mark.parametrize("mode", ["forward", "deferred"])
@pytest.mark.parametrize("names", SETS, ids=lambda n: n[-1].split("/")[0])
def test_asset_frame(ctx, scene_dir, names, mode):
    label = f"{names[-1].split('/')[0]} {mode}"
    p = A.scene(os.path.join(scene_dir, "assets", f"{names[-1].split('/')[0]}_{mode}.scene"), names, mode=mode,
                width=640, height=400, ssao="on" if mode == "deferred" else "off")
    _check(ctx, p, label, 0.05)


def test_readme_diablo_frame(ctx, scene_dir):
    """The README's console-output run (README.md:190-223) as the reference
    builds it by default, and as the scene loads here without filter keys:
    Diablo alone, forward Blinn-Phong, 1280x800, shadow on with
    SOFT_SHADOW_PCSS drawing its samples from the random stream
    (shadow.h:15-16, shadow.cpp:65-106) for every fragment that passes the
    depth test, SSAA and SSAO off.  Depth and shadow map bit-exact, the
    stream consumed to the same draw, image within 1 LSB."""
    p = A.scene(os.path.join(scene_dir, "assets", "readme_diablo.scene"), ["diablo_pose"])
    sc = _check(ctx, p, "README Diablo 1280x800 forward, reference stream", 0.1)
    I = sc.info
    m, _ = sc.model(0)
    assert (I.width, I.height, I.deferred, I.shadow_on, I.shadow.mode, len(m["vi"])) == \
        (1280, 800, 0, 1, F.SHADOW_REF_PCSS, 5022)
    assert ctx.rng_draws() == O.rng_draws(), (ctx.rng_draws(), O.rng_draws())


def test_readme_diablo_fixed_disk(ctx, scene_dir):
    """The same frame with the reference's PCSS sample counts and sizes on
    one fixed disk (`shadowfilter pcss 32 64 2.5`)."""
    p = A.scene(os.path.join(scene_dir, "assets", "readme_diablo_fixed.scene"), ["diablo_pose"],
                extra=("shadowfilter pcss 32 64 2.5",))
    sc = _check(ctx, p, "README Diablo 1280x800 forward, fixed disk", 0.1)
    assert (sc.info.deferred, sc.info.shadow.mode) == (0, F.SHADOW_PCSS)


def test_diablo_deferred_ssaa(ctx, scene_dir):
    """The deferred path (G-buffer, SSAO + blur, lighting) with SSAA 2x
    (internal 1280x800) on Diablo over the plane."""
    p = A.scene(os.path.join(scene_dir, "assets", "diablo_ssaa.scene"), ["plane", "diablo_pose"], mode="deferred",
                width=640, height=400, ssao="on", ssaa="on 2")
    sc = _check(ctx, p, "Diablo deferred SSAA 2x", 0.3)
    assert (sc.info.ssaa_on, sc.info.ssaa_k) == (1, 2)
