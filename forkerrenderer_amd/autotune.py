"""Raster autotuner: picks the screen-tile edge (8, 16 or 32 pixels; the
workgroup shape follows: 64 lanes for 8x8, 256 lanes for 16x16 and 32x32)
separately for the shadow pass and the main pass by timing whole frames of
the scene on the device.  Results never depend on the choice (each pixel's
depth resolve is a minimum over (depth, submission id) keys, the same for
any tile shape or list order); only speed does."""
import itertools
import time

from . import PASS_GEOMETRY, PASS_SHADOW, lib

TILES = (8, 16, 32)


def _frame_ms(ctx, scene, reps):
    ctx.render_scene(scene)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.render_scene(scene)
    ctx.sync()
    return (time.perf_counter() - t0) * 1e3 / reps


def autotune(ctx, scene, tiles=TILES, reps=3):
    """Returns ({"shadow": t, "main": t}, table of ms per combination) and
    leaves the best combination set on the context."""
    table = {}
    for ts, tm in itertools.product(tiles, tiles):
        ctx._chk(lib().fgl_set_pass_tile_size(ctx._h, PASS_SHADOW, ts), "tile")
        ctx._chk(lib().fgl_set_pass_tile_size(ctx._h, PASS_GEOMETRY, tm), "tile")
        table[(ts, tm)] = _frame_ms(ctx, scene, reps)
    best = min(table, key=table.get)
    ctx._chk(lib().fgl_set_pass_tile_size(ctx._h, PASS_SHADOW, best[0]), "tile")
    ctx._chk(lib().fgl_set_pass_tile_size(ctx._h, PASS_GEOMETRY, best[1]), "tile")
    return {"shadow": best[0], "main": best[1]}, {f"{k[0]}/{k[1]}": v for k, v in table.items()}
