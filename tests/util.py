"""Helpers shared by the parity tests: run a scene through the GPU path and
through the oracle, and compare buffers."""
import numpy as np

import forkerrenderer_amd as F
import oracle as O


def scene_arrays(scene):
    I = scene.info
    models = [scene.model(i) for i in range(I.n_models)]
    mats = [scene.material(i) for i in range(I.n_materials)]
    texs = [scene.texture(i) for i in range(I.n_textures)]
    return models, mats, texs


def gpu_frame(ctx, scene, tile=16):
    ctx.set_tile_size(tile)
    ctx.render_scene(scene)
    I = scene.info
    out = {"image": ctx.read(F.TARGET_IMAGE), "depth": ctx.read(F.TARGET_DEPTH),
           "fb": ctx.read(F.TARGET_FRAME).reshape(3, -1)}
    if I.shadow_on:
        out["shadow"] = ctx.read(F.TARGET_SHADOW)
        out["shadow_depth"] = ctx.read(F.TARGET_SHADOW_DEPTH)
    if I.deferred:
        out["gbuf"] = ctx.read(F.TARGET_GBUFFER).reshape(F.GBUF_PLANES, -1)
        out["ao"] = ctx.read(F.TARGET_AO)
    return out


def oracle_frame(scene):
    models, mats, texs = scene_arrays(scene)
    return O.render_frame(models, mats, texs, scene.info)


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def assert_bitexact(name, got, want):
    g, w = bits(got), bits(want)
    assert g.shape == w.shape, f"{name}: shape {g.shape} vs {w.shape}"
    bad = np.nonzero(g != w)[0]
    assert len(bad) == 0, (f"{name}: {len(bad)} of {g.size} differ; first at {bad[:5]}: "
                           f"got {np.asarray(got).reshape(-1)[bad[:5]]} want {np.asarray(want).reshape(-1)[bad[:5]]}")


def lsb_stats(got, want):
    d = np.abs(got.astype(np.int32) - want.astype(np.int32))
    return int(d.max()) if d.size else 0, float((d <= 1).mean()) if d.size else 1.0
