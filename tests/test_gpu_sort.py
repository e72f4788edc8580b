"""The binning's device radix sort (fgl_sort_pairs) against numpy's stable
argsort: keys and values bit-exact, equal keys in input order.  Sizes around
the 2048-item block (1, 2047, 2048, 2049, ragged multi-block counts), key
widths of 1-32 bits, uniform keys, all-equal keys, and clustered runs like
the (tile, triangle) pairs of a mesh in submission order (few digits per
block, the case whose scatter writes short runs)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _check(ctx, keys, vals, bits):
    print(f"sort n={keys.size} bits={bits}", flush=True)
    gk, gv = ctx.sort_pairs(keys, vals, bits)
    o = np.argsort(keys, kind="stable")
    assert np.array_equal(gk, keys[o])
    assert np.array_equal(gv, vals[o])


@pytest.mark.parametrize("n", [1, 2, 100, 2047, 2048, 2049, 100_003, 3_000_017])
@pytest.mark.parametrize("bits", [1, 8, 15, 32])
def test_uniform_keys(ctx, n, bits):
    rng = np.random.default_rng(n * 40 + bits)
    hi = 1 << bits
    keys = rng.integers(0, hi, n, dtype=np.uint64).astype(np.uint32)
    vals = np.arange(n, dtype=np.uint32)
    _check(ctx, keys, vals, bits)


def test_equal_keys_keep_their_order(ctx):
    n = 50_001
    keys = np.full(n, 1234, np.uint32)
    vals = np.random.default_rng(1).integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
    _check(ctx, keys, vals, 15)


@pytest.mark.parametrize("n", [500_000, 4_000_037])
def test_clustered_tile_runs(ctx, n):
    """Tile ids of a 240 x 135 grid visited in small screen-space patches with
    repeats, as a dense mesh's pairs come out of the emit."""
    rng = np.random.default_rng(n)
    cx = np.cumsum(rng.integers(-1, 2, n // 50 + 1)) % 240
    cy = np.cumsum(rng.integers(-1, 2, n // 50 + 1)) % 135
    patch = np.repeat(cy * 240 + cx, 50)[:n]
    jitter = rng.integers(0, 3, n) + 240 * rng.integers(0, 3, n)
    keys = ((patch + jitter) % (240 * 135)).astype(np.uint32)
    vals = np.arange(n, dtype=np.uint32)
    _check(ctx, keys, vals, 15)


def test_rejects_bad_arguments(ctx):
    import forkerrenderer_amd as F
    with pytest.raises(F.FglError):
        ctx.sort_pairs(np.zeros(4, np.uint32), np.zeros(4, np.uint32), 0)
    k, v = ctx.sort_pairs(np.zeros(0, np.uint32), np.zeros(0, np.uint32), 8)
    assert k.size == 0 and v.size == 0
