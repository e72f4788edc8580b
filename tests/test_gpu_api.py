"""GPU checks of the ABI's state semantics against the oracle: draws split
over several flushes, lazily applied clears, reads before any pass, tile-size
independence, and the general (non-integer-exact) coverage path for vertices
behind the camera."""
import os

import numpy as np
import pytest

import forkerrenderer_amd as F
from forkerrenderer_amd import scenegen as G
import oracle as O
from tests.util import assert_bitexact, gpu_frame, oracle_frame

pytestmark = pytest.mark.gpu


def test_reads_before_any_pass_return_clear_values(ctx):
    ctx.init_targets(40, 30)
    d = ctx.read(F.TARGET_DEPTH)
    g = ctx.read(F.TARGET_GBUFFER)
    ao = ctx.read(F.TARGET_AO)
    fb = ctx.read(F.TARGET_FRAME)
    assert np.all(d == np.finfo(np.float32).max) and np.all(g == 0) and np.all(ao == 1) and np.all(fb == 0)
    ctx.clear_color(0.12, 0.25, 0.5)
    fb = ctx.read(F.TARGET_FRAME).reshape(3, -1)
    assert np.all(fb[0] == np.float32(0.12)) and np.all(fb[2] == np.float32(0.5))


def test_draws_split_over_flushes_equal_one_pass(ctx, scene_dir):
    """Geometry pass drawn as two models in one batch, and again with a read
    (forcing a flush) between the draws: identical targets, equal to the oracle."""
    p = G.tiny_scene(os.path.join(scene_dir, "split"), seed=21, width=80, height=60, subdiv=3, tex=16)
    sc = F.Scene(p)
    want = oracle_frame(sc)
    gpu_frame(ctx, sc)  # sets the context's matrices, lights, shadow map exactly as the scene does
    ids = ctx.upload_scene(sc)
    I = sc.info
    outs = []
    for split in (False, True):
        ctx.init_targets(I.width, I.height)
        ctx.begin_pass(F.PASS_GEOMETRY)
        ctx.draw_model(*ids[0])
        if split:
            ctx.read(F.TARGET_DEPTH)
        ctx.draw_model(*ids[1])
        ctx.end_pass()
        outs.append((ctx.read(F.TARGET_DEPTH), ctx.read(F.TARGET_GBUFFER)))
    assert_bitexact("depth split", outs[1][0], outs[0][0])
    assert_bitexact("gbuf split", outs[1][1], outs[0][1])
    assert_bitexact("depth vs oracle", outs[0][0], want["depth"])
    assert_bitexact("gbuf vs oracle", outs[0][1], want["gbuf"].reshape(-1))


def test_camera_inside_geometry_general_path(ctx, scene_dir):
    """Camera close to the sphere: some triangles have vertices behind the
    camera (w <= 0), window coordinates far outside the exact integer range,
    and INT_MIN truncations.  Coverage/depth/G-buffer stay bit-exact."""
    d = os.path.join(scene_dir, "inside")
    p = G.tiny_scene(d, seed=22, width=72, height=56, subdiv=3, tex=16)
    txt = open(p).read().replace("camera persp 0.3 1.0 2.6 0 -0.2 0", "camera persp 0.2 0.1 0.75 0 -0.3 -2")
    open(p, "w").write(txt)
    sc = F.Scene(p)
    for tile in (8, 32):
        got = gpu_frame(ctx, sc, tile)
        want = oracle_frame(sc)
        assert (want["depth"] < 3e38).sum() > 200
        assert_bitexact("depth", got["depth"], want["depth"])
        assert_bitexact("gbuf", got["gbuf"], want["gbuf"])


def test_ssao_reference_blur_and_multilight(ctx, scene_dir):
    p = G.tiny_scene(os.path.join(scene_dir, "ml"), seed=23, width=90, height=60, subdiv=3, tex=16, ssao=True,
                     shadow="pcss", extra="aoblur reference\nmultilight on\nlight dir -0.3 -1 -0.2 0.5 0.5 0.6")
    sc = F.Scene(p)
    assert sc.info.n_lights == 2 and sc.info.ssao.blur == 1
    got, want = gpu_frame(ctx, sc), oracle_frame(sc)
    assert_bitexact("ao", got["ao"], want["ao"])
    d = np.abs(got["image"].astype(int) - want["image"].astype(int))
    assert (d <= 1).mean() >= 0.999


def test_forward_config1_full_size(ctx, scene_dir):
    """Config 1 at its real size: 640x400 forward Blinn-Phong, hard shadow,
    repeat + bilinear noise texture."""
    p = G.config1(os.path.join(scene_dir, "c1full"))
    sc = F.Scene(p)
    got, want = gpu_frame(ctx, sc), oracle_frame(sc)
    assert_bitexact("depth", got["depth"], want["depth"])
    assert_bitexact("shadow", got["shadow"], want["shadow"])
    d = np.abs(got["image"].astype(int) - want["image"].astype(int))
    print("c1 image: max diff", d.max(), "fraction within 1 LSB", (d <= 1).mean())
    assert (d <= 1).mean() >= 0.999


def test_nan_initial_depth_follows_the_submission_order_fold(ctx, scene_dir):
    """The reference's fold `replace when !(d >= best)` lets any fragment
    replace a NaN depth -- not a minimum.  Depth written with scattered NaNs,
    then a geometry pass over it: covered pixels equal the fresh render
    (and the oracle), uncovered ones keep what was written."""
    d = os.path.join(scene_dir, "nan")
    p = G.tiny_scene(d, seed=24, width=64, height=48, subdiv=2, tex=16)
    sc = F.Scene(p)
    I = sc.info
    W, H = I.width, I.height
    want = oracle_frame(sc)
    fresh = gpu_frame(ctx, sc)          # also leaves the scene's matrices and lights set
    assert_bitexact("fresh depth", fresh["depth"], want["depth"])
    ids = ctx.upload_scene(sc)
    init = np.full(W * H, np.finfo(np.float32).max, np.float32)
    init[::7] = np.nan
    covered = want["depth"] < 3e38
    assert (covered & np.isnan(init)).sum() > 50
    for tile in (8, 16, 32):
        ctx.set_tile_size(tile)
        ctx.init_targets(W, H)
        ctx.write(F.TARGET_DEPTH, init)
        ctx.begin_pass(F.PASS_GEOMETRY)
        for mid, M in ids:
            ctx.draw_model(mid, M)
        ctx.end_pass()
        depth = ctx.read(F.TARGET_DEPTH)
        gbuf = ctx.read(F.TARGET_GBUFFER).reshape(F.GBUF_PLANES, -1)
        assert_bitexact(f"covered depth, tile {tile}", depth[covered], want["depth"][covered])
        assert_bitexact(f"uncovered depth, tile {tile}", depth[~covered], init[~covered])
        assert_bitexact(f"covered gbuf, tile {tile}", gbuf[:, covered], fresh["gbuf"][:, covered])


def _nan_aware_equal(name, got, want):
    gn, wn = np.isnan(got), np.isnan(want)
    assert np.array_equal(gn, wn), f"{name}: NaN positions differ ({gn.sum()} vs {wn.sum()})"
    assert_bitexact(name, got[~gn], want[~wn])


def test_nan_fragments_follow_the_submission_order_fold(ctx):
    """A vertex at NaN truncates to INT_MIN window coordinates with a NaN
    depth, so its triangles cover real pixels with NaN fragments; an infinite
    depth gives NaN where its weight is zero.  A NaN fragment always replaces
    the stored depth and whatever follows it replaces the NaN: the winner is
    the depth minimum over the fragments after the pixel's last NaN fragment.
    Shadow pass with identity matrices, every tile size, against the oracle's
    in-order fold (NaN positions equal, every other value bit-exact)."""
    rng = np.random.default_rng(8)
    W, H = 64, 48
    n = 60
    pos = rng.uniform(-1.2, 1.2, (n, 3)).astype(np.float32)
    pos[:, 2] = rng.uniform(-0.9, 0.9, n)
    pos[3] = np.nan
    pos[7, 2] = np.inf
    faces = rng.integers(0, n, (80, 3)).astype(np.int32)
    md = {"pos": pos, "uv": np.zeros((n, 2), np.float32), "nrm": np.tile([[0, 0, 1]], (n, 1)).astype(np.float32),
          "tan": None, "vi": faces, "ti": faces, "ni": faces, "face_material": np.zeros(len(faces), np.int32),
          "support_pbr": False}
    eye = np.eye(4, dtype=np.float32)
    sc = O.OracleScene([(md, eye)], [O.Material()], [])
    vp = O.make_viewport(0, 0, W, H)
    fr = O.make_frame(eye, eye, eye, vp, vp, [0, 0, 0], 1, None, [])
    want_d, want_s = O.shadow_pass(sc, fr, W, H)
    assert np.isnan(want_d).sum() > 50 and (want_d < 3e38).sum() > 1000
    m = F.Material()
    for f in ("map_kd", "map_ks", "map_ke", "map_bump", "map_base", "map_rough", "map_metal", "map_ao", "map_pbr_ke"):
        setattr(m, f, -1)
    mat = ctx.upload_material(m)
    md["face_material"] = np.full(len(faces), mat, np.int32)
    mid = ctx.upload_model(md)
    ctx.set_viewport(0, 0, W, H, shadow=True)
    ctx.set_matrix(F.MAT_LIGHT_SPACE, eye)
    for tile in (8, 16, 32):
        ctx.set_pass_tile_size(F.PASS_SHADOW, tile)
        ctx.init_shadow_buffer(W, H)
        ctx.begin_pass(F.PASS_SHADOW)
        ctx.draw_model(mid, eye)
        ctx.end_pass()
        _nan_aware_equal(f"shadow depth, tile {tile}", ctx.read(F.TARGET_SHADOW_DEPTH), want_d)
        _nan_aware_equal(f"shadow map, tile {tile}", ctx.read(F.TARGET_SHADOW), want_s)


def test_pair_count_overflow_is_an_error(ctx):
    """140,000 triangles that each cover a 3840 x 2160 target give 1.8e10
    (tile, triangle) pairs at tile 8, more than the 32-bit pair offsets hold:
    the pass stops with FGL_ERR_CAPACITY before any pair is written, and the
    context renders normally afterwards."""
    n = 140_000
    tri = np.array([[-3, -3, 0.5], [3, -3, 0.5], [0, 3, 0.5]], np.float32)
    pos = np.tile(tri, (n, 1))
    faces = np.arange(3 * n, dtype=np.int32).reshape(n, 3)
    m = F.Material()
    for f in ("map_kd", "map_ks", "map_ke", "map_bump", "map_base", "map_rough", "map_metal", "map_ao", "map_pbr_ke"):
        setattr(m, f, -1)
    mat = ctx.upload_material(m)
    md = {"pos": pos, "uv": np.zeros((3 * n, 2), np.float32), "nrm": np.tile([[0, 0, 1]], (3 * n, 1)).astype(np.float32),
          "tan": None, "vi": faces, "ti": faces, "ni": faces, "face_material": np.full(n, mat, np.int32),
          "support_pbr": False}
    mid = ctx.upload_model(md)
    eye = np.eye(4, dtype=np.float32)
    W, H = 3840, 2160
    ctx.set_viewport(0, 0, W, H, shadow=True)
    ctx.set_matrix(F.MAT_LIGHT_SPACE, eye)
    ctx.set_pass_tile_size(F.PASS_SHADOW, 8)
    ctx.init_shadow_buffer(W, H)
    ctx.begin_pass(F.PASS_SHADOW)
    ctx.draw_model(mid, eye)
    print("pair overflow pass", flush=True)
    with pytest.raises(F.FglError, match="2\\^32"):
        ctx.end_pass()
    ctx.set_pass_tile_size(F.PASS_SHADOW, 16)
    # a pass of one of those triangles works and covers every texel
    md1 = dict(md, pos=tri, uv=np.zeros((3, 2), np.float32), nrm=np.tile([[0, 0, 1]], (3, 1)).astype(np.float32),
               vi=faces[:1], ti=faces[:1], ni=faces[:1], face_material=np.full(1, mat, np.int32))
    one = ctx.upload_model(md1)
    ctx.init_shadow_buffer(W, H)
    ctx.begin_pass(F.PASS_SHADOW)
    ctx.draw_model(one, eye)
    ctx.end_pass()
    d = ctx.read(F.TARGET_SHADOW_DEPTH)
    assert np.all(np.abs(d - np.float32(0.75)) < 1e-6), "every texel covered at window z 0.75"
