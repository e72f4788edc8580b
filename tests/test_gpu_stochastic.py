"""The reference's random stream on the device against the oracle: the
generator's unit-sphere and unit-disk vectors and draw counts (bit-exact,
across calls that end mid-block), Render::DoSSAO with the stream (AO
bit-exact), and whole deferred and forward frames with the reference's PCF /
PCSS filters drawing from the stream (G-buffer, AO and depth bit-exact, the
stream consumed to the same draw, image within the usual +-1 LSB)."""
import os

import numpy as np
import pytest

import forkerrenderer_amd as F
from forkerrenderer_amd import scenegen as G
import oracle as O
from tests.test_oracle_rng import mt_raw, stream_disk, stream_sphere
from tests.util import assert_bitexact, gpu_frame, lsb_stats, oracle_frame

pytestmark = pytest.mark.gpu


def test_device_stream_matches_the_oracle(ctx):
    ctx.rng_seed(5489)
    sph = ctx.rng_vectors(3, 20000)
    raw = mt_raw(400000)
    want, used = stream_sphere(raw, 20000)
    assert_bitexact("sphere", sph, want)
    assert ctx.rng_draws() == used
    dsk = ctx.rng_vectors(2, 30000)
    want_d, used_d = stream_disk(raw[used:], 30000)
    assert_bitexact("disk", dsk, want_d)
    assert ctx.rng_draws() == used + used_d
    # short requests that end inside a 624-word block, then a long one
    ctx.rng_seed(77)
    O.rng_seed(77)
    for kind, n in ((3, 1), (2, 1), (3, 7), (2, 5), (3, 150), (2, 333), (3, 5000), (2, 1), (3, 2)):
        got = ctx.rng_vectors(kind, n)
        want = O.rng_sphere(n) if kind == 3 else O.rng_disk(n)
        assert_bitexact(f"kind {kind} n {n}", got, want)
        assert ctx.rng_draws() == O.rng_draws(), (kind, n)
    ctx.rng_seed(5489)


@pytest.mark.parametrize("seed", [5489, 77])
def test_device_stream_in_parallel_chunks_matches_numpy(ctx, seed):
    """Requests of more than one chunk (256 blocks, 159,744 draws, while 256
    such chunks cover the request) run as parallel chunks, each reached by
    jumping ahead from the context's array:
    vectors bit-exact with numpy's MT19937 through the reference's rejection
    loops, draw counts equal, and the stream continues exactly from the
    array the commit rebuilds (small requests before, between and after,
    so the big ones start and end inside a block)."""
    ctx.rng_seed(seed)
    raw = mt_raw(26_000_000, seed=seed)
    at = 0
    for kind, n in ((3, 7), (2, 3_000_000), (2, 5), (3, 2_000_000), (2, 1), (3, 1)):
        print(f"seed {seed}: kind {kind} n {n}", flush=True)
        got = ctx.rng_vectors(kind, n)
        want, used = (stream_sphere if kind == 3 else stream_disk)(raw[at:], n)
        assert_bitexact(f"kind {kind} n {n}", got, want)
        at += used
        assert ctx.rng_draws() == at, (kind, n, ctx.rng_draws(), at)
    ctx.rng_seed(5489)


def test_device_stream_long_chunks_match_numpy(ctx):
    """A request of more than 256 chunks of 256 blocks (17 M disk vectors,
    ~43 M draws) runs as chunks of 512 blocks, jumping with their own
    polynomials: vectors bit-exact with numpy's MT19937, the draw count equal,
    and the stream continuing exactly after it."""
    ctx.rng_seed(5489)
    raw = mt_raw(44_000_000)
    at = 0
    for kind, n in ((2, 3), (2, 17_000_000), (3, 5)):
        print(f"kind {kind} n {n}", flush=True)
        got = ctx.rng_vectors(kind, n)
        want, used = (stream_sphere if kind == 3 else stream_disk)(raw[at:], n)
        assert_bitexact(f"kind {kind} n {n}", got, want)
        at += used
        assert ctx.rng_draws() == at, (kind, n, ctx.rng_draws(), at)
    assert at > 256 * 256 * 624, at
    ctx.rng_seed(5489)


def _deferred_targets(ctx, sc):
    """Geometry pass of a scene through the context (the frame's matrices and
    lights set as fgl_render_scene sets them); the frame's stream starts at
    the reference's default seed."""
    ctx.rng_seed(5489)
    gpu_frame(ctx, sc)
    return ctx.read(F.TARGET_DEPTH), ctx.read(F.TARGET_GBUFFER).reshape(F.GBUF_PLANES, -1)


def test_reference_ssao_matches_the_oracle(ctx, scene_dir):
    p = G.tiny_scene(os.path.join(scene_dir, "rssao"), seed=31, width=64, height=48, subdiv=3, tex=16,
                     ssao=True, extra="ssaokernel reference\naoblur none")
    sc = F.Scene(p)
    I = sc.info
    W, H = I.width, I.height
    depth, g = _deferred_targets(ctx, sc)
    want = oracle_frame(sc)
    assert_bitexact("depth", depth, want["depth"])
    assert_bitexact("gbuf", g, want["gbuf"])
    # the frame's own SSAO ran with the stream from 5489: AO equals the oracle's
    ao_frame = ctx.read(F.TARGET_AO)
    assert_bitexact("frame ao", ao_frame, want["ao"])
    assert len(np.unique(want["ao"])) > 5, "AO too uniform to test"
    # again on the same targets, the stream continued by one call
    fr = O.make_frame(np.asarray(ctx_view(sc)), proj(sc), np.eye(4), O.make_viewport(0, 0, W, H),
                      O.make_viewport(0, 0, W, H), I.eye, 0, None, [])
    ctx.rng_seed(1234)
    O.rng_seed(1234)
    cfg = F.SsaoCfg.from_buffer_copy(I.ssao)
    cfg.blur = 0
    for rep in range(2):
        ctx.ssao(cfg)
        got = ctx.read(F.TARGET_AO)
        ref = O.ssao_reference(fr, W, H, depth, g)
        assert_bitexact(f"ao call {rep}", got, ref)
        assert ctx.rng_draws() == O.rng_draws()


def ctx_view(sc):
    I = sc.info
    return O.make_lookat(list(I.eye), list(I.look_at))


def proj(sc):
    I = sc.info
    ratio = np.float32(I.width) / np.float32(I.height)
    return O.make_perspective(45.0, ratio, 0.01, 20.0)


@pytest.mark.parametrize("filt", ["refpcf", "refpcss"])
def test_stochastic_frames_match_the_oracle(ctx, scene_dir, filt):
    p = G.tiny_scene(os.path.join(scene_dir, "r" + filt), seed=32, width=64, height=48, subdiv=3, tex=16,
                     ssao=True, extra=f"ssaokernel reference\naoblur reference\nshadowfilter {filt}")
    sc = F.Scene(p)
    assert sc.info.shadow.mode == (F.SHADOW_REF_PCF if filt == "refpcf" else F.SHADOW_REF_PCSS)
    print(f"stochastic frame {filt}", flush=True)
    ctx.rng_seed(5489)
    got = gpu_frame(ctx, sc)
    device_draws = ctx.rng_draws()
    want = oracle_frame(sc)
    assert device_draws == O.rng_draws()
    assert_bitexact("shadow", got["shadow"], want["shadow"])
    assert_bitexact("gbuf", got["gbuf"], want["gbuf"])
    assert_bitexact("ao", got["ao"], want["ao"])
    mx, frac = lsb_stats(got["image"], want["image"])
    assert frac >= 0.999, (mx, frac)
    # the filter's samples were consumed after SSAO's
    P = sc.info.width * sc.info.height
    assert device_draws > 32 * P * 3


@pytest.mark.parametrize("tile", [8, 16, 32])
@pytest.mark.parametrize("filt", ["refpcf", "refpcss"])
def test_forward_stochastic_frames_match_the_oracle(ctx, scene_dir, filt, tile):
    """Forward pass with the reference's PCF / PCSS drawing from the stream:
    the shaders draw for every fragment that passes the depth test when it is
    drawn, overdrawn ones included, in raster order (triangles in order, px
    outer, py inner); the device replays that sequence in each tile.  Depth
    and shadow map bit-exact, the stream consumed to the same draw, image
    within 1 LSB; every tile size (each has its own walk kernel)."""
    p = G.tiny_scene(os.path.join(scene_dir, "f" + filt), seed=33, mode="forward", width=96, height=72, subdiv=3,
                     tex=16, extra=f"shadowfilter {filt}")
    sc = F.Scene(p)
    assert not sc.info.deferred
    print(f"forward stochastic {filt} tile {tile}", flush=True)
    got = gpu_frame(ctx, sc, tile)
    device_draws = ctx.rng_draws()
    want = oracle_frame(sc)
    assert device_draws == O.rng_draws(), (device_draws, O.rng_draws())
    assert_bitexact("depth", got["depth"], want["depth"])
    assert_bitexact("shadow", got["shadow"], want["shadow"])
    mx, frac = lsb_stats(got["image"], want["image"])
    print(f"  draws {device_draws}, image max diff {mx}, within 1 LSB {frac:.6f}", flush=True)
    assert frac >= 0.999, (mx, frac)
    # every shaded pixel drew at least one sample set (2 draws per try)
    covered = int((want["depth"] < 3e38).sum())
    assert device_draws > 2 * 32 * covered


@pytest.mark.parametrize("pbr", [False, True], ids=["phong", "pbr"])
def test_forward_stochastic_ssaa_and_two_lights(ctx, scene_dir, pbr):
    """The forward pass's fragment sequence at SSAA 2x (the internal frame's
    pixels) with a point and a directional light (only the point light is
    shadowed and draws samples), Blinn-Phong and Cook-Torrance."""
    p = G.tiny_scene(os.path.join(scene_dir, f"fssaa{int(pbr)}"), seed=35, mode="forward", width=60, height=44,
                     subdiv=3, tex=16, ssaa=2, pbr=pbr,
                     extra="shadowfilter refpcss\nmultilight on\nlight dir -0.3 -1 -0.2 0.5 0.5 0.6")
    sc = F.Scene(p)
    assert sc.info.ssaa_on == 1 and sc.info.n_lights == 2
    got = gpu_frame(ctx, sc)
    device_draws = ctx.rng_draws()
    want = oracle_frame(sc)
    assert device_draws == O.rng_draws(), (device_draws, O.rng_draws())
    assert_bitexact("depth", got["depth"], want["depth"])
    assert_bitexact("shadow", got["shadow"], want["shadow"])
    mx, frac = lsb_stats(got["image"], want["image"])
    assert frac >= 0.999, (mx, frac)


def test_forward_stochastic_passes_split_over_flushes_equal_one_pass(ctx, scene_dir):
    """The depth buffer and the stream carry over between forward passes as
    in the reference: the models drawn in one pass, or in two passes (the
    second starting from the first's depth), give the same depth, frame and
    stream position, equal to the scene render."""
    p = G.tiny_scene(os.path.join(scene_dir, "fsplit"), seed=34, mode="forward", width=80, height=60, subdiv=3,
                     tex=16, extra="shadowfilter refpcss")
    sc = F.Scene(p)
    full = gpu_frame(ctx, sc)   # sets the scene's matrices, lights and shadow map
    full_draws = ctx.rng_draws()
    ids = ctx.upload_scene(sc)
    assert len(ids) >= 2
    I = sc.info
    outs = []
    for split in (False, True):
        ctx.init_targets(I.width, I.height)
        ctx.clear_color(0.12, 0.12, 0.12)
        ctx.rng_seed(5489)
        ctx.begin_pass(F.PASS_FORWARD)
        for j, (mid, M) in enumerate(ids):
            if split and j == 1:
                ctx.end_pass()
                ctx.begin_pass(F.PASS_FORWARD)
            ctx.draw_model(mid, M)
        ctx.end_pass()
        outs.append((ctx.read(F.TARGET_DEPTH), ctx.read(F.TARGET_FRAME), ctx.rng_draws()))
    for name, k in (("depth", 0), ("frame", 1)):
        assert_bitexact(f"{name} split", outs[1][k], outs[0][k])
    assert outs[0][2] == outs[1][2] == full_draws
    assert_bitexact("depth vs scene", outs[0][0], full["depth"])
    assert_bitexact("frame vs scene", outs[0][1], full["fb"].reshape(-1))


def test_batched_views_refuse_the_stochastic_filters(ctx, scene_dir):
    """Batched views take the deterministic filters only: a view with
    FGL_SHADOW_REF_* is refused when the batch is made (no device work is
    started), and the context still works afterwards."""
    from forkerrenderer_amd import views as VW
    v = VW.uneven_views(ctx, n_views=1, seed=3, subdivs=(2,), min_px=32 * 32, max_px=64 * 64, shadow_size=64,
                        tex=16)[0]
    assert v.desc.shadow_on
    v.desc.shadow.mode = F.SHADOW_REF_PCSS
    with pytest.raises(F.FglError, match="deterministic shadow filters"):
        VW.Batch(ctx, [v])
    with pytest.raises(F.FglError):
        ctx.rng_vectors(4, 10)
    p2 = G.tiny_scene(os.path.join(scene_dir, "rok"), seed=33)
    sc2 = F.Scene(p2)
    got = gpu_frame(ctx, sc2)
    want = oracle_frame(sc2)
    assert_bitexact("depth after errors", got["depth"], want["depth"])
