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


def _batch_ms(batch, reps):
    batch.render()
    batch.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        batch.render()
    batch.sync()
    return (time.perf_counter() - t0) * 1e3 / reps


def autotune_batch(batch, tiles=TILES, occupancies=(2, 4, 8), margins=(1.0, 1.25, 2.0), reps=2):
    """Batched views: the raster tile edge of the shadow and main passes (the
    rasteriser's workgroup shape follows the tile), the persistent
    rasteriser workgroups per compute unit, then the (tile, triangle) pair
    capacity of each pass ('bin capacity') as the measured pair count times
    a margin.  Times graph replays of the whole batch; leaves the best
    setting on the batch.  Returns (setting, table of ms per setting)."""
    table = {}
    for ts, tm in itertools.product(tiles, tiles):
        for occ in occupancies:
            batch.set_tiles(ts, tm)
            batch.set_occupancy(occ)
            batch.set_capacity(0, 0)
            table[f"{ts}/{tm}/{occ}/auto"] = _batch_ms(batch, reps)
    best = min(table, key=table.get)
    ts, tm, occ = (int(x) for x in best.split("/")[:3])
    batch.set_tiles(ts, tm)
    batch.set_occupancy(occ)
    batch.set_capacity(0, 0)
    batch.render()
    batch.sync()
    st = batch.stats()
    for m in margins:
        cap = (max(1024, int(st["pairs_shadow"] * m)), max(1024, int(st["pairs_main"] * m)))
        batch.set_capacity(*cap)
        table[f"{ts}/{tm}/{occ}/x{m}"] = _batch_ms(batch, reps)
    best = min((k for k in table if k.startswith(f"{ts}/{tm}/{occ}/")), key=table.get)
    m = best.split("/")[3]
    if m == "auto":
        batch.set_capacity(0, 0)
    else:
        f = float(m[1:])
        batch.set_capacity(max(1024, int(st["pairs_shadow"] * f)), max(1024, int(st["pairs_main"] * f)))
    batch.render()
    batch.sync()
    return {"shadow_tile": ts, "main_tile": tm, "workgroups_per_cu": occ, "pair_capacity": m}, table
