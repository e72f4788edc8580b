"""The benchmark configurations at their real sizes, against the oracle.

Config 2: 1280x800 deferred PBR, 327,682 triangles, 2048^2 PBR maps,
2048^2 shadow map with 5x5 PCF, 64-sample kernel SSAO + sigma=2 blur --
the whole frame.

Config 3: 1920x1080 with SSAA 2x (3840x2160 internal), 5,242,882 triangles,
PCSS 16/32 area 0.05, kernel SSAO + sigma=2 blur, a point and a directional
light.  Shadow map, shadow-pass depth, depth and all 19 G-buffer planes are
compared bit for bit over the whole frame; the unblurred SSAO on row bands
(the oracle's SSAO takes ~25 s for all 2160 rows on one core); the blur over
the whole frame (oracle blur of the device's unblurred AO); lighting and
quantisation on row bands (oracle lighting of the device's G-buffer and AO,
both already checked), and the SSAA resolve of the whole image from the
device's quantised frame by the integer closed form.

Bars: bit-exact for coverage, depth, shadow map, G-buffer and AO; 8-bit
output within +-1 LSB on >= 99.9 % of the pixels.
"""
import os

import numpy as np
import pytest

import forkerrenderer_amd as F
import oracle as O
from forkerrenderer_amd import scenegen as G
from forkerrenderer_amd import views as VW
from tests.util import assert_bitexact, lsb_stats, oracle_frame, scene_arrays

pytestmark = pytest.mark.gpu


def _rows(a, W, y0, y1):
    return a.reshape(-1, W)[y0:y1] if a.ndim == 1 else a.reshape(a.shape[0], -1, W)[:, y0:y1]


def test_config2_whole_frame(ctx, scene_dir):
    p = G.config2(os.path.join(scene_dir, "c2"))
    sc = F.Scene(p)
    I = sc.info
    assert (I.width, I.height, I.shadow_map_size, I.ssao.n_kernel) == (1280, 800, 2048, 64)
    print("config 2: render on the device", flush=True)
    ctx.set_tile_size(16)
    ctx.render_scene(sc)
    got = {"depth": ctx.read(F.TARGET_DEPTH), "gbuf": ctx.read(F.TARGET_GBUFFER).reshape(F.GBUF_PLANES, -1),
           "ao": ctx.read(F.TARGET_AO), "shadow": ctx.read(F.TARGET_SHADOW),
           "shadow_depth": ctx.read(F.TARGET_SHADOW_DEPTH), "image": ctx.read(F.TARGET_IMAGE)}
    print("config 2: oracle", flush=True)
    want = oracle_frame(sc)
    ntri = sum(len(sc.model(i)[0]["vi"]) for i in range(I.n_models))
    assert ntri == 327682
    assert (want["depth"] < 3e38).mean() > 0.2     # sphere and ground cover ~26 % of the frame
    for k in ("shadow", "shadow_depth", "depth", "gbuf", "ao"):
        assert_bitexact(k, got[k], want[k])
    mx, frac = lsb_stats(got["image"], want["image"])
    print(f"config 2 image: max diff {mx}, within 1 LSB {frac:.6f}")
    assert frac >= 0.999, (mx, frac)


def test_config3_full_size(ctx, scene_dir):
    p = G.config3(os.path.join(scene_dir, "c3"))
    sc = F.Scene(p)
    I = sc.info
    models, mats, texs = scene_arrays(sc)
    assert sum(len(m["vi"]) for m, _ in models) == 5242882
    st = O.frame_setup(I)
    W, H, SW, SH, k = st["W"], st["H"], st["SW"], st["SH"], st["k"]
    assert (W, H, SW, SH, k) == (3840, 2160, 2048, 2048, 2)
    print("config 3: render on the device", flush=True)
    ctx.set_pass_tile_size(F.PASS_SHADOW, 16)
    ctx.set_pass_tile_size(F.PASS_GEOMETRY, 16)
    ctx.render_scene(sc)
    g_depth = ctx.read(F.TARGET_DEPTH)
    g_gbuf = ctx.read(F.TARGET_GBUFFER).reshape(F.GBUF_PLANES, -1)
    g_shadow, g_sdepth = ctx.read(F.TARGET_SHADOW), ctx.read(F.TARGET_SHADOW_DEPTH)
    g_ao = ctx.read(F.TARGET_AO)
    g_fb = ctx.read(F.TARGET_FRAME).reshape(3, -1)
    g_img = ctx.read(F.TARGET_IMAGE)
    # the same frame's SSAO without the blur (depth and G-buffer are unchanged)
    cfg = F.SsaoCfg()
    F.C.memmove(F.C.byref(cfg), F.C.byref(I.ssao), F.C.sizeof(F.SsaoCfg))
    cfg.blur = 0
    ctx.ssao(cfg)
    g_raw = ctx.read(F.TARGET_AO)

    osc = O.OracleScene(models, mats, texs)
    print("config 3: oracle shadow pass", flush=True)
    sdepth, smap = O.shadow_pass(osc, st["fr_shadow"], SW, SH)
    assert_bitexact("shadow_depth", g_sdepth, sdepth)
    assert_bitexact("shadow", g_shadow, smap)
    print("config 3: oracle geometry pass", flush=True)
    depth, gbuf = O.geometry_pass(osc, st["fr"], W, H)
    cover = depth < 3e38
    assert cover.mean() > 0.5
    assert_bitexact("depth", g_depth, depth)
    assert_bitexact("gbuf", g_gbuf, gbuf)
    del gbuf
    fr = st["fr"]
    lib = O.lib()
    bands = [(0, 48), (1040, 1136), (2112, 2160)]   # bottom rows, middle, top rows
    R = I.ssao.blur_radius
    try:
        for y0, y1 in bands:
            print(f"config 3: oracle SSAO and lighting, rows {y0}-{y1}", flush=True)
            # unblurred AO on the band widened by the blur radius: every blurred
            # value the band's lighting reads is then checked
            a0, a1 = max(0, y0 - R), min(H, y1 + R)
            lib.orc_set_rows(a0, a1)
            raw = O.ssao_kernel(fr, W, H, depth, g_gbuf, I.ssao)
            assert_bitexact(f"raw AO rows {a0}-{a1}", _rows(g_raw, W, a0, a1), _rows(raw, W, a0, a1))
            lib.orc_set_rows(y0, y1)
            fb = O.lighting_pass(fr, W, H, g_gbuf, g_ao, smap, SW, SH)
            q = O.quantize_rgb(_rows(fb, W, y0, y1).reshape(3, -1), W, y1 - y0)
            gq = O.quantize_rgb(_rows(g_fb, W, y0, y1).reshape(3, -1), W, y1 - y0)
            mx, frac = lsb_stats(gq, q)
            print(f"config 3 rows {y0}-{y1}: max diff {mx}, within 1 LSB {frac:.6f}", flush=True)
            assert frac >= 0.999, (y0, y1, mx, frac)
    finally:
        lib.orc_set_rows(0, 1 << 30)
    occluded = (_rows(g_raw, W, 1040, 1136) < 1)[_rows(cover, W, 1040, 1136)]
    assert occluded.mean() > 0.05, "middle band has too little occlusion to test"
    assert_bitexact("blurred AO", g_ao, O.blur_separable(g_raw, W, H, I.ssao.blur_weights, I.ssao.blur_radius))
    assert np.array_equal(g_img, O.ssaa_resolve(O.quantize_rgb(g_fb, W, H), W, H, k)), "SSAA resolve"


def test_config4_neural_ao_full_frame(ctx, scene_dir):
    """Config 4: the config-3 scene at 1920x1080 without SSAA, neural AO in
    place of SSAO + blur, with the shipped trained weights.  Depth and
    G-buffer bit-exact against the oracle; the device's down pass, network
    and up pass against the network oracle (block inputs, BF16 forward,
    bilinear up pass) over the whole frame -- max 2e-2, mean 2e-4 absolute
    (FP32 accumulation order inside the MFMA tiles, expf); and the frame's
    AO error against plain SSAO + blur, bounded to catch regressions of the
    trained network."""
    from forkerrenderer_amd import nao as NA
    from oracle import nao_ref as R
    from tests.test_gpu_nao import bf16_bits
    p = G.config4(os.path.join(scene_dir, "c4"))
    sc = F.Scene(p)
    I = sc.info
    st = O.frame_setup(I)
    W, H = st["W"], st["H"]
    assert (W, H) == (1920, 1080)
    ctx.render_scene(sc)
    depth = ctx.read(F.TARGET_DEPTH)
    gbuf = ctx.read(F.TARGET_GBUFFER).reshape(F.GBUF_PLANES, -1)
    plain = ctx.read(F.TARGET_AO).reshape(H, W)
    ck = NA.load_checkpoint(NA.DEFAULT_WEIGHTS)
    params = ck["state"][:NA.param_count()]
    net = NA.NeuralAO(ctx, batch=1, crop=64, params=params)
    net.infer()
    pred = ctx.read(F.TARGET_AO).reshape(H, W)
    net.close()
    print("config 4: oracle geometry pass", flush=True)
    models, mats, texs = scene_arrays(sc)
    od, og = O.geometry_pass(O.OracleScene(models, mats, texs), st["fr"], W, H)
    assert_bitexact("depth", depth, od)
    assert_bitexact("gbuf", gbuf, og)
    # network oracle on the whole network grid
    g3 = gbuf.reshape(F.GBUF_PLANES, H, W)
    cov = R.gbuffer_coverage(depth.reshape(H, W), g3)
    S = int(ck["scale"])
    x, lcov = R.block_inputs_gbuf(depth.reshape(H, W), g3, st["view"].astype(np.float32), S)
    _, y = R.forward(params, R.bf16_bits_to_float(bf16_bits(x))[None], bf16=True)
    want = R.upsample(np.where(lcov, y[0], 1.0).astype(np.float32), cov, S)
    err = np.abs(pred - want)[cov]
    print(f"config 4 network vs oracle: max {err.max():.2e} mean {err.mean():.2e}", flush=True)
    assert err.max() < 2e-2 and err.mean() < 2e-4
    assert np.all(pred[~cov] == 1.0)
    mae = float(np.abs(pred - plain)[cov].mean())
    print(f"config 4 neural AO vs plain SSAO + blur: MAE {mae:.4f}", flush=True)
    assert mae < 0.035, mae


def _oracle_view(spec, host):
    """One view of the batched-views configuration through the oracle, set
    up as fgl_render_view does: the shadow pass from the point light (ortho
    +-3 x aspect, 0.1..20, looking at the origin) on the S x S map, the
    geometry pass (perspective 45 degrees, 0.01..20)."""
    W, H, S = spec.width, spec.height, spec.shadow_size
    ratio = np.float32(W) / np.float32(H)
    light = spec.light
    Lm = O.mat_mul(O.make_ortho(np.float32(-3.0) * ratio, np.float32(3.0) * ratio, -3.0, 3.0, 0.1, 20.0),
                   O.make_lookat(list(light[1]), (0, 0, 0)))
    V = O.make_lookat(list(spec.eye), list(spec.look_at))
    P = O.make_perspective(45.0, ratio, 0.01, 20.0)
    vp, svp = O.make_viewport(0, 0, W, H), O.make_viewport(0, 0, S, S)
    cfg = VW.pcf_grid(5)
    lights = [(light[0], list(light[1]), list(light[2]))]
    fr_s = O.make_frame(np.eye(4), np.eye(4), Lm, vp, svp, spec.eye, 1, cfg, lights)
    fr = O.make_frame(V, P, Lm, vp, svp, spec.eye, 1, cfg, lights)
    Mg, Ms = spec.matrices()
    texs = [{"bytes": t, "wrap": F.WRAP_REPEAT, "filter": F.FILTER_LINEAR} for t in host["maps"]]
    osc = O.OracleScene([(host["ground"], Mg), (host["sphere"], Ms)], [VW.pbr_material_desc([0, 1, 2, 3])], texs)
    sdepth, smap = O.shadow_pass(osc, fr_s, S, S)
    depth, g = O.geometry_pass(osc, fr, W, H)
    return dict(fr=fr, V=V, sdepth=sdepth, smap=smap, depth=depth, gbuf=g)


@pytest.mark.timeout(900)
def test_config5_batched_views(ctx):
    """Config 5: 64 views of distinct seeded scenes of the 1280x800 generator
    (2048^2 maps, 2048^2 shadow maps with 5x5 PCF, subdivision 4-8), sizes
    log-uniform 256^2 .. 1920x1080, neural AO with the shipped weights, one
    batch replayed from its HIP graph.  Every target of every view equals
    the view rendered alone (fgl_render_view) bit for bit; the four smallest
    views against the oracle: shadow map, depth and G-buffer bit-exact,
    neural AO against the network oracle (max 2e-2, mean 2e-4), image
    within +-1 LSB (oracle lighting of the device's G-buffer and AO)."""
    from forkerrenderer_amd import nao as NA
    from oracle import nao_ref as R
    from tests.test_gpu_nao import bf16_bits
    specs = VW.view_specs()
    assert len(specs) == 64 and {s.subdiv for s in specs} == {4, 5, 6, 7, 8}
    print("config 5: generating and uploading 64 scenes", flush=True)
    views = VW.distinct_views(ctx, specs)
    ck = NA.load_checkpoint(NA.DEFAULT_WEIGHTS)
    params = ck["state"][:NA.param_count()]
    net = NA.NeuralAO(ctx, batch=1, crop=64, params=params)
    b = VW.Batch(ctx, views, VW.AO_NEURAL, net)
    b.render()
    assert b.sync() == 0
    b.render()                      # graph replay
    assert b.sync() == 0
    st = b.stats()
    print(f"config 5: {st['pixels']} pixels, {st['tris_main']} triangles per pass", flush=True)
    targets = [F.TARGET_DEPTH, F.TARGET_GBUFFER, F.TARGET_AO, F.TARGET_FRAME, F.TARGET_IMAGE, F.TARGET_SHADOW,
               F.TARGET_SHADOW_DEPTH]
    for i, v in enumerate(views):
        VW.render_view(ctx, v, VW.AO_NEURAL, net)
        for t in targets:
            got, want = b.read(i, t), ctx.read(t)
            assert np.array_equal(got.view(np.uint8), np.ascontiguousarray(want).view(np.uint8)), \
                f"view {i} ({v.desc.width}x{v.desc.height}) target {t}"
    print("config 5: all 64 views equal their single-view renders", flush=True)
    small = sorted(range(64), key=lambda i: specs[i].width * specs[i].height)[:4]
    for i in small:
        spec = specs[i]
        W, H = spec.width, spec.height
        print(f"config 5: view {i} ({W}x{H}, subdivision {spec.subdiv}) against the oracle", flush=True)
        o = _oracle_view(spec, VW.scene_host(spec))
        assert_bitexact(f"view {i} shadow", b.read(i, F.TARGET_SHADOW), o["smap"])
        assert_bitexact(f"view {i} shadow depth", b.read(i, F.TARGET_SHADOW_DEPTH), o["sdepth"])
        depth = b.read(i, F.TARGET_DEPTH)
        gbuf = b.read(i, F.TARGET_GBUFFER).reshape(F.GBUF_PLANES, -1)
        assert_bitexact(f"view {i} depth", depth, o["depth"])
        assert_bitexact(f"view {i} gbuf", gbuf, o["gbuf"])
        ao = b.read(i, F.TARGET_AO).reshape(H, W)
        g3 = gbuf.reshape(F.GBUF_PLANES, H, W)
        cov = R.gbuffer_coverage(depth.reshape(H, W), g3)
        x, lcov = R.block_inputs_gbuf(depth.reshape(H, W), g3, o["V"].astype(np.float32), NA.SCALE)
        _, y = R.forward(params, R.bf16_bits_to_float(bf16_bits(x))[None], bf16=True)
        want_ao = R.upsample(np.where(lcov, y[0], 1.0).astype(np.float32), cov, NA.SCALE)
        err = np.abs(ao - want_ao)[cov]
        assert err.max() < 2e-2 and err.mean() < 2e-4, (i, err.max(), err.mean())
        fb = O.lighting_pass(o["fr"], W, H, gbuf, ao.reshape(-1), o["smap"], spec.shadow_size, spec.shadow_size)
        mx, frac = lsb_stats(b.read(i, F.TARGET_IMAGE), O.quantize_rgb(fb, W, H))
        assert frac >= 0.999, (i, mx, frac)
    b.close()
    net.close()
