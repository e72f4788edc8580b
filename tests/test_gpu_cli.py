"""The command-line program (the reference's src/main.cpp:18-55 with
Output::* of src/output.cpp:12-86) against the oracle: it renders a scene
file on the GPU and writes the reference's output images; each image is
decoded with the oracle's TGA reader and compared with the oracle frame's
buffers quantised by the oracle (Buffer3f/Buffer1f::GenerateImage,
src/buffer.cpp:20-32,113-126).  Depth, shadow map, G-buffer and AO images
are byte-identical (their floats are bit-exact); the shaded images are
within +-1 LSB on >= 99.9 % of the bytes."""
import os
import subprocess

import numpy as np
import pytest

import forkerrenderer_amd as F
from forkerrenderer_amd import scenegen as G
import oracle as O
from oracle.objload import tga_read
from tests.util import oracle_frame

EXE = os.path.join(os.path.dirname(F.__file__), "_lib", "ForkerRendererAMD")


def _run_cli(scene, cwd):
    print(f"cli {scene}", flush=True)
    r = subprocess.run([EXE, scene], cwd=cwd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    return r.stdout


def _img(path, W, H, ch):
    with open(path, "rb") as fh:
        raw = fh.read()
    assert raw[2] == (11 if ch == 1 else 10) and raw[16] == 8 * ch, "RLE grey / RLE true colour, bits per pixel"
    a = tga_read(path)[::-1]          # rows back to buffer order (row 0 = bottom)
    assert a.shape[:2] == (H, W)
    return np.ascontiguousarray(a[..., :ch]).reshape(-1)


def _lsb(got, want):
    d = np.abs(got.astype(np.int32) - want.astype(np.int32))
    return int(d.max()), float((d <= 1).mean())


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["deferred", "forward"])
def test_cli_writes_the_reference_outputs(scene_dir, mode):
    d = os.path.join(scene_dir, "cli_" + mode)
    p = G.tiny_scene(d, seed=41, width=56, height=40, mode=mode, ssao=mode == "deferred", ssaa=2, shadow="pcss")
    sc = F.Scene(p)
    I = sc.info
    out_dir = os.path.join(d, "output")
    os.makedirs(out_dir, exist_ok=True)
    for f in os.listdir(out_dir):
        os.remove(os.path.join(out_dir, f))
    log = _run_cli(p, d)
    assert "Output TGA File" in log
    want = oracle_frame(sc)
    W, H = I.width * 2, I.height * 2
    S = I.shadow_map_size
    files = sorted(os.listdir(out_dir))
    expect = ["framebuffer.tga", "framebuffer_SSAA.tga", "shadowmap.tga", "zbuffer.tga"]
    if I.deferred:
        expect += ["gbuffer_albedo.tga", "gbuffer_ambient_occlusion.tga", "gbuffer_normal.tga",
                   "gbuffer_param.tga", "gbuffer_shading_type.tga", "gbuffer_worldpos.tga"]
    assert files == sorted(expect)
    o = lambda n: os.path.join(out_dir, n)
    assert np.array_equal(_img(o("zbuffer.tga"), W, H, 1), O.quantize_grey(want["depth"], W, H))
    assert np.array_equal(_img(o("shadowmap.tga"), S, S, 1), O.quantize_grey(want["shadow"], S, S))
    if I.deferred:
        g = want["gbuf"]
        for name, plane in (("normal", 0), ("worldpos", 3), ("albedo", 9), ("param", 15)):
            assert np.array_equal(_img(o(f"gbuffer_{name}.tga"), W, H, 3),
                                  O.quantize_rgb(g[plane:plane + 3], W, H)), name
        assert np.array_equal(_img(o("gbuffer_shading_type.tga"), W, H, 1), O.quantize_grey(g[18], W, H))
        ao = O.quantize_grey(want["ao"], W, H)
        assert np.array_equal(_img(o("gbuffer_ambient_occlusion.tga"), W, H, 1), ao)
        assert len(np.unique(ao)) > 3, "AO image too uniform to test"
    mx, frac = _lsb(_img(o("framebuffer.tga"), W, H, 3), O.quantize_rgb(want["fb"], W, H))
    assert frac >= 0.999, (mx, frac)
    mx, frac = _lsb(_img(o("framebuffer_SSAA.tga"), I.width, I.height, 3), want["image"])
    assert frac >= 0.999, (mx, frac)


def test_cli_usage_and_missing_scene(scene_dir):
    """Argument and load errors end the program before any device work."""
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "Usage" in r.stderr
    r = subprocess.run([EXE, os.path.join(scene_dir, "no_such.scene")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "[error]" in r.stderr
