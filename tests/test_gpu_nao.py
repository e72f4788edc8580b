"""Neural AO on the GPU against the CPU network oracle (oracle/nao_ref.py):
the down pass (block inputs) and staged training crops, forward through the
up pass, gradients, the Adam step, deterministic training, checkpoint
resume, and the audit fallback.  The network runs on the 4x coarser grid
(512 x 384 frame -> 128 x 96 network pixels)."""
import numpy as np
import pytest

import forkerrenderer_amd as F
from forkerrenderer_amd import nao as NA
from oracle import nao_ref as R

pytestmark = pytest.mark.gpu

W, H, CROP, BATCH = 512, 384, 64, 4
S = NA.SCALE
LW, LH = W // S, H // S
XY = [(0, 0), (64, 32), (31, 17), (10, 30)]    # crop origins on the network grid


@pytest.fixture(scope="module")
def world(ctx, tmp_path_factory):
    sset = NA.SceneSet(ctx, str(tmp_path_factory.mktemp("nao")), subdiv=4, tex_size=256)
    cfg = NA.ssao_cfg()
    NA.render_frame(ctx, sset, 5003, W, H, cfg)
    frame = {"depth": ctx.read(F.TARGET_DEPTH).reshape(H, W),
             "gbuf": ctx.read(F.TARGET_GBUFFER).reshape(F.GBUF_PLANES, H, W),
             "ao": ctx.read(F.TARGET_AO).reshape(H, W),
             "view": NA.make_lookat(*NA.camera(5003)).astype(np.float32)}
    assert (frame["depth"] < 3e38).mean() > 0.3, "test frame should be mostly covered"
    return sset, cfg, frame


def _params(seed=7):
    p = R.init_params(seed)
    rng = np.random.default_rng(seed)
    offs = np.cumsum([0] + [s for _, s in R.SIZES])
    for k, (name, s) in enumerate(R.SIZES):
        if name.startswith("b"):
            p[offs[k]:offs[k + 1]] += rng.normal(0, 0.05, s).astype(np.float32)
    return p.astype(np.float32)


def net_inputs(frame):
    """The network grid's inputs and coverage (oracle down pass)."""
    cov = R.gbuffer_coverage(frame["depth"], frame["gbuf"])
    x, lcov = R.block_inputs_gbuf(frame["depth"], frame["gbuf"], frame["view"], S)
    return x, lcov, cov


def test_nonfinite_gbuffer_pixels_are_uncovered(ctx, world):
    """A NaN normal or position in the G-buffer makes the pixel uncovered: it
    takes no part in its block's inputs or target (the staged crop equals
    the oracle's block inputs with the pixel uncovered), the gradients stay
    finite, and its AO at inference is 1."""
    sset, cfg, frame = world
    NA.render_frame(ctx, sset, 5003, W, H, cfg)
    g = ctx.read(F.TARGET_GBUFFER).reshape(F.GBUF_PLANES, H, W).copy()
    cover = frame["depth"] < 3e38
    ys, xs = np.nonzero(cover)
    picks = [(int(ys[k]), int(xs[k])) for k in (0, len(ys) // 2, len(ys) - 1)]
    g[0, picks[0][0], picks[0][1]] = np.nan      # normal x
    g[4, picks[1][0], picks[1][1]] = np.inf      # position y
    g[2, picks[2][0], picks[2][1]] = np.nan      # normal z
    ctx.write(F.TARGET_GBUFFER, g)
    bad = dict(frame)
    bad["gbuf"] = g
    x, lcov, cov = net_inputs(bad)
    for y, xx in picks:
        assert not cov[y, xx]
    n = NA.NeuralAO(ctx, batch=1, crop=64, params=_params())
    n.infer()
    ao = ctx.read(F.TARGET_AO).reshape(H, W)
    for y, xx in picks:
        assert ao[y, xx] == 1.0
        X, Y = xx // S, y // S
        x0, y0 = min(max(X - 32, 0), LW - 64), min(max(Y - 32, 0), LH - 64)
        n.stage(0, [(x0, y0)])
        feat, _, mask = n.read_batch()
        assert np.array_equal(feat[0], bf16_bits(x[y0:y0 + 64, x0:x0 + 64]))
        assert np.array_equal(mask[0].astype(bool), R.crop_loss_mask(lcov[y0:y0 + 64, x0:x0 + 64], x0, y0, 64, LW, LH))
    g_grad, st = n.gradients()
    assert np.all(np.isfinite(g_grad))
    n.close()


def bf16_bits(x):
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    return (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF).astype(np.uint16)


@pytest.fixture(scope="module")
def staged(ctx, world):
    sset, cfg, frame = world
    net = NA.NeuralAO(ctx, batch=BATCH, crop=CROP, params=_params())
    NA.render_frame(ctx, sset, 5003, W, H, cfg)
    net.stage(0, XY)
    feat, tgt, mask = net.read_batch()
    yield net, feat, tgt, mask
    net.close()


def test_staged_crops_match_gbuffer(world, staged):
    """Crops of the network grid: inputs (block means of the pixel features
    and the covered fraction, in BF16), targets (block means of the plain
    AO over covered pixels) and mask, bit for bit against the oracle."""
    _, _, frame = world
    _, feat, tgt, mask = staged
    x, lcov, cov = net_inputs(frame)
    t, _ = R.block_targets(frame["ao"], cov, S)
    refb = bf16_bits(x)
    assert 0.005 < (x[..., 7][lcov] < 1).mean() < 0.5, "blocks on silhouettes should be partly covered"
    for i, (cx, cy) in enumerate(XY):
        sl = (slice(cy, cy + CROP), slice(cx, cx + CROP))
        assert np.array_equal(feat[i], refb[sl]), f"crop {i}"
        assert np.array_equal(tgt[i].view(np.uint32), t[sl].view(np.uint32)), f"crop {i} target"
        assert np.array_equal(mask[i].astype(bool), R.crop_loss_mask(lcov[sl], cx, cy, CROP, LW, LH))


def test_gradients_match_oracle(staged):
    net, feat, tgt, mask = staged
    p = _params()
    net.set_params(p)
    g, st = net.gradients()
    featf = R.bf16_bits_to_float(feat)
    loss, gr, (sse, sae, cnt) = R.loss_and_grads(p, featf, tgt, mask, bf16=True)
    assert st[2] == cnt
    assert abs(st[0] - sse) <= 1e-4 * sse + 1e-6
    assert abs(st[1] - sae) <= 1e-4 * sae + 1e-6
    offs = np.cumsum([0] + [s for _, s in R.SIZES])
    for k, (name, _) in enumerate(R.SIZES):
        a, b = g[offs[k]:offs[k + 1]].astype(np.float64), gr[offs[k]:offs[k + 1]]
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        # FP32 (device) vs FP64 (oracle) sums can move a BF16 rounding by one
        # unit (2^-8 relative) at a few activations / data gradients.
        assert rel < 5e-3, (name, rel)


def test_forward_matches_oracle(ctx, world, staged):
    sset, cfg, frame = world
    net = staged[0]
    p = _params()
    net.set_params(p)
    NA.render_frame(ctx, sset, 5003, W, H, cfg)
    net.infer()
    y = ctx.read(F.TARGET_AO).reshape(H, W)
    x, lcov, cov = net_inputs(frame)
    featf = R.bf16_bits_to_float(bf16_bits(x)).reshape(1, LH, LW, 16)
    _, yr = R.forward(p, featf, bf16=True)
    low = np.where(lcov, yr[0], 1.0).astype(np.float32)
    want = R.upsample(low, cov, S)
    assert np.all(y[~cov] == 1.0)
    err = np.abs(y - want)[cov]
    assert err.max() < 2e-2 and err.mean() < 2e-4, (err.max(), err.mean())


def test_adam_step_matches_numpy(staged):
    net = staged[0]
    p = _params()
    net.set_params(p)
    g, _ = net.gradients()
    lr = np.float32(1e-3)
    net.train_step(float(lr))
    p1 = net.params()
    f = np.float32
    m = f(0.9) * f(0) + (f(1) - f(0.9)) * g
    v = f(0.999) * f(0) + (f(1) - f(0.999)) * g * g
    c1, c2 = f(1.0 - 0.9), f(1.0 - 0.999)
    want = p - lr * (m / c1) / (np.sqrt(v / c2) + f(1e-8))
    assert np.array_equal(p1.view(np.uint32), want.astype(np.float32).view(np.uint32))


def test_training_is_deterministic_and_resumes_bitwise(ctx, world):
    sset, cfg, _ = world
    NA.render_frame(ctx, sset, 5003, W, H, cfg)

    def fresh():
        n = NA.NeuralAO(ctx, batch=BATCH, crop=CROP, params=_params())
        n.stage(0, XY)
        return n
    a = fresh()
    for _ in range(4):
        a.train_step(1e-3)
    b = fresh()
    for _ in range(2):
        b.train_step(1e-3)
    st, step = b.state()
    b.close()
    c = fresh()
    c.load_state(st, step)
    for _ in range(2):
        c.train_step(1e-3)
    pa, sa = a.state()
    pc, sc = c.state()
    assert sa == sc == 4
    assert np.array_equal(pa.view(np.uint32), pc.view(np.uint32))
    a.close()
    c.close()


def test_training_reduces_loss(ctx, world):
    sset, cfg, _ = world
    NA.render_frame(ctx, sset, 5003, W, H, cfg)
    n = NA.NeuralAO(ctx, batch=BATCH, crop=CROP, seed=1)
    n.stage(0, XY)
    first = n.train_step(1e-3)
    for _ in range(60):
        last = n.train_step(1e-3)
    n.close()
    assert last[0] < 0.5 * first[0], (first, last)


def test_audit_falls_back_on_bad_weights(ctx, world):
    sset, cfg, _ = world
    p = R.init_params(0)
    p[-1] = -30.0   # head bias: AO ~ 0 everywhere
    n = NA.NeuralAO(ctx, batch=1, crop=64, params=p)
    NA.render_frame(ctx, sset, 5003, W, H, cfg)
    plain = ctx.read(F.TARGET_AO).reshape(H, W)
    cover = ctx.read(F.TARGET_DEPTH).reshape(H, W) < 3e38
    n.infer()
    pred = ctx.read(F.TARGET_AO).reshape(H, W)
    n.close()
    fixed, audited, failed = NA.audit_fallback(pred, plain, cover, 0, every=1)
    assert audited > 0 and failed >= 0.9 * audited, (audited, failed)
    assert np.abs(fixed - plain)[cover].mean() < np.abs(pred - plain)[cover].mean()


@pytest.mark.parametrize("size", [(512, 384), (333, 201), (40, 24), (1920, 1080)],
                         ids=lambda s: f"{s[0]}x{s[1]}")
def test_fused_inference_equals_layered(ctx, world, size):
    """The two fused inference kernels (layers 1-2; layers 3-4 + head on the
    6 x 6 phase grids, halos recomputed in LDS) give the same bits as one
    kernel per layer, on grids that split into whole, ragged and partial
    tiles (network grids 128x96, 84x51, 10x6, 480x270)."""
    sset, cfg, _ = world
    w, h = size
    net = NA.NeuralAO(ctx, batch=1, crop=64, params=_params(11))
    try:
        NA.render_frame(ctx, sset, 5011, w, h, cfg)
        cov = ctx.read(F.TARGET_DEPTH) < 3e38
        assert cov.any()
        out = {}
        for fused in (True, False):
            net.set_fused(fused)
            net.infer()
            out[fused] = ctx.read(F.TARGET_AO).copy()
        assert np.array_equal(out[True].view(np.uint32), out[False].view(np.uint32)), \
            int((out[True] != out[False]).sum())
        assert np.all(out[True][~cov] == 1.0) and (out[True][cov] < 1.0).any()
    finally:
        net.close()


def test_network_ao_stays_on_its_grid_until_read(ctx, world):
    """fgl_nao_infer leaves the AO on the network grid: the deferred lighting
    pass interpolates it there, and reading the AO target runs the up pass.
    Lighting from the grid and from the written-out plane give the same
    frame bit for bit; an SSAO pass replaces the pending network AO; closing
    the network writes its pending AO out first."""
    sset, cfg, frame = world
    net = NA.NeuralAO(ctx, batch=1, crop=64, params=_params())
    NA.render_frame(ctx, sset, 5003, W, H, cfg)
    plain = ctx.read(F.TARGET_AO).copy()
    cur = {"depth": ctx.read(F.TARGET_DEPTH).reshape(H, W),
           "gbuf": ctx.read(F.TARGET_GBUFFER).reshape(F.GBUF_PLANES, H, W), "view": frame["view"]}
    net.infer()
    ctx.draw_screen_space_pixels()                       # AO from the grid
    fb_grid = ctx.read(F.TARGET_FRAME).copy()
    ao = ctx.read(F.TARGET_AO).copy()                    # the up pass runs here
    ctx.draw_screen_space_pixels()                       # AO from the plane
    fb_plane = ctx.read(F.TARGET_FRAME).copy()
    assert np.array_equal(fb_grid.view(np.uint32), fb_plane.view(np.uint32))
    x, lcov, c2 = net_inputs(cur)
    cov = c2.reshape(-1)
    assert np.all(ao[~cov] == 1.0) and not np.array_equal(ao, plain)
    # the oracle's network and up pass on this frame give the same AO
    featf = R.bf16_bits_to_float(bf16_bits(x)).reshape(1, LH, LW, 16)
    _, yr = R.forward(_params(), featf, bf16=True)
    want = R.upsample(np.where(lcov, yr[0], 1.0).astype(np.float32), c2, S).reshape(-1)
    assert np.abs(ao - want)[cov].max() < 2e-2
    # SSAO replaces the pending network AO
    net.infer()
    ctx.ssao(cfg)
    assert np.array_equal(ctx.read(F.TARGET_AO).view(np.uint32), plain.view(np.uint32))
    # a network closed with its AO pending writes it out first
    net.infer()
    net.close()
    assert np.array_equal(ctx.read(F.TARGET_AO).view(np.uint32), ao.view(np.uint32))
