"""Batched uneven views: every target of every view of a batch equals the
single-view render of that view bit for bit -- across tile sizes,
rasteriser occupancy, graph replays and a forced pair-capacity overflow."""
import numpy as np
import pytest

import forkerrenderer_amd as F
from forkerrenderer_amd import nao as NA
from forkerrenderer_amd import views as VW

pytestmark = pytest.mark.gpu

TARGETS = [F.TARGET_DEPTH, F.TARGET_GBUFFER, F.TARGET_FRAME, F.TARGET_IMAGE, F.TARGET_SHADOW,
           F.TARGET_SHADOW_DEPTH]


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint32) if a.dtype == np.float32 else a


@pytest.fixture(scope="module")
def world(ctx):
    views = VW.uneven_views(ctx, n_views=7, seed=11, subdivs=(2, 3, 4), min_px=48 * 48, max_px=320 * 240,
                            shadow_size=256, tex=64)
    # one view without shadows, one with an extra model instance
    v = views[3].desc
    views[3] = VW.View(v.width, v.height, tuple(v.eye), tuple(v.look_at),
                       list(zip(views[3].models.tolist(), views[3].mats.reshape(-1, 4, 4))),
                       [(0, (2.0, 5.0, 5.0), (3.0, 3.0, 3.0))], shadow=None)
    v = views[5]
    extra = list(zip(v.models.tolist(), v.mats.reshape(-1, 4, 4)))
    extra.append((int(v.models[1]), F.make_model_matrix((1.2, 0.2, -0.5), 30.0, 0.4)))
    d = v.desc
    views[5] = VW.View(d.width, d.height, tuple(d.eye), tuple(d.look_at), extra,
                       [(0, (2.0, 5.0, 5.0), (3.0, 3.0, 3.0))], VW.pcf_grid(5), 256)
    net = NA.NeuralAO(ctx, batch=1, crop=64, params=NA.init_params(3))
    singles = []
    for i, view in enumerate(views):
        VW.render_view(ctx, view, VW.AO_NEURAL, net)
        out = {t: ctx.read(t).copy() for t in TARGETS if view.desc.shadow_on or t not in
               (F.TARGET_SHADOW, F.TARGET_SHADOW_DEPTH)}
        out[F.TARGET_AO] = ctx.read(F.TARGET_AO).copy()
        singles.append(out)
    yield views, net, singles
    net.close()


def check(batch, views, singles, label):
    for i, ref in enumerate(singles):
        for t, want in ref.items():
            got = batch.read(i, t)
            assert np.array_equal(bits(got), bits(want)), f"{label}: view {i} target {t}"


def test_batch_equals_single_views(ctx, world):
    views, net, singles = world
    b = VW.Batch(ctx, views, VW.AO_NEURAL, net)
    b.render()
    assert b.sync() == 0
    check(b, views, singles, "eager")
    st = b.stats()
    assert st["views"] == len(views) and st["pairs_main"] > 0 and st["pairs_main"] <= st["cap_main"]
    b.close()


@pytest.mark.parametrize("tiles,occ", [((8, 8), 1), ((32, 16), 2), ((16, 32), 8)])
def test_batch_graph_replays_across_tiles(ctx, world, tiles, occ):
    views, net, singles = world
    b = VW.Batch(ctx, views, VW.AO_NEURAL, net)
    b.set_tiles(*tiles)
    b.set_occupancy(occ)
    for rep in range(3):
        b.render()
        assert b.sync() == 0
        check(b, views, singles, f"tiles {tiles} occ {occ} replay {rep}")
        ctx.sync()   # eager work on the same stream between replays
    b.close()


def test_batch_overflow_rerenders(ctx, world):
    views, net, singles = world
    b = VW.Batch(ctx, views, VW.AO_NEURAL, net)
    b.set_capacity(200, 300)
    b.render()
    assert b.sync() >= 1
    check(b, views, singles, "after overflow")
    st = b.stats()
    assert st["cap_main"] >= st["pairs_main"] and st["cap_shadow"] >= st["pairs_shadow"]
    b.render()
    assert b.sync() == 0
    check(b, views, singles, "graph after overflow")
    b.close()


def test_batch_without_ao(ctx, world):
    views, _, _ = world
    ref = []
    for view in views[:3]:
        VW.render_view(ctx, view, VW.AO_NONE)
        ref.append({t: ctx.read(t).copy() for t in (F.TARGET_DEPTH, F.TARGET_FRAME, F.TARGET_IMAGE)})
    b = VW.Batch(ctx, views[:3], VW.AO_NONE)
    b.render()
    b.sync()
    b.render()
    b.sync()
    check(b, views[:3], ref, "no AO")
    b.close()


def test_batch_after_tables_and_ao_buffers_move(ctx, world, scene_dir):
    """Resources the batch's captured graph addresses can move after the
    batch was planned: the context's material and texture tables when new
    uploads outgrow them, the neural-AO activations when a single-frame
    inference runs on a frame larger than the whole batch.  Replays after
    each move still equal the single-view renders."""
    import os
    from forkerrenderer_amd import scenegen as G
    views, net, singles = world
    b = VW.Batch(ctx, views, VW.AO_NEURAL, net)
    b.render()
    assert b.sync() == 0
    b.render()            # the graph is captured here
    assert b.sync() == 0
    check(b, views, singles, "before")
    # many new materials and textures: both tables are reallocated
    tex = np.zeros((4, 4, 3), np.uint8)
    m = F.Material()
    for f in ("map_kd", "map_ks", "map_ke", "map_bump", "map_base", "map_rough", "map_metal", "map_ao", "map_pbr_ke"):
        setattr(m, f, -1)
    for _ in range(300):
        ctx.upload_texture(tex)
        ctx.upload_material(m)
    for rep in range(2):
        b.render()
        assert b.sync() == 0
        check(b, views, singles, f"after table growth, replay {rep}")
    # a single-frame neural AO larger than the batch moves the activations
    px = b.stats()["pixels"]
    W, H = 1024, 1 + px // 1024
    p = G.tiny_scene(os.path.join(scene_dir, "batch_big_nao"), seed=3, width=W, height=H, subdiv=2, tex=16)
    ctx.render_scene(F.Scene(p))
    net.infer()
    ctx.sync()
    for rep in range(2):
        b.render()
        assert b.sync() == 0
        check(b, views, singles, f"after activation move, replay {rep}")
    b.close()
