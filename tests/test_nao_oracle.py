"""Pins the neural-AO oracle: the backward formulas against central finite
differences of the float64 network, and the BF16 rounding against values the
format fixes."""
import numpy as np

from oracle import nao_ref as R


def test_bf16_rounding_fixed_points():
    vals = np.array([1.0, -2.5, 0.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, 3.140625], np.float32)
    want = np.array([1.0, -2.5, 0.0, 1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -6, 3.140625])
    # 1 + 2^-8 is a tie -> even (1.0); 1 + 3*2^-8 is a tie -> even (1 + 2^-6)
    assert np.array_equal(R.bf16_round(vals), want)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    N, H, W = 1, 19, 21   # larger than the widest dilated tap reach
    feat = np.zeros((N, H, W, 16))
    feat[..., :7] = rng.normal(0, 1, (N, H, W, 7))
    target = rng.uniform(0, 1, (N, H, W))
    mask = rng.uniform(0, 1, (N, H, W)) < 0.8
    p = R.init_params(3).astype(np.float64)
    offs = np.cumsum([0] + [s for _, s in R.SIZES])
    for k, (name, s) in enumerate(R.SIZES):   # nonzero biases exercise the bias gradients
        if name.startswith("b"):
            p[offs[k]:offs[k + 1]] = rng.normal(0, 0.1, s)
    p[R.NPARAM - 1] = 0.3
    loss, g, _ = R.loss_and_grads(p, feat, target, mask, bf16=False)
    idx = []
    for k in range(len(R.SIZES)):
        lo, hi = offs[k], offs[k + 1]
        cand = np.arange(lo, hi)
        if R.SIZES[k][0] == "W1":
            cand = cand[(cand - lo) % 16 < 7]   # padding channels have zero input
        idx += list(rng.choice(cand, size=min(4, len(cand)), replace=False))
    eps = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        lp, _, _ = R.loss_and_grads(pp, feat, target, mask, bf16=False)
        lm, _, _ = R.loss_and_grads(pm, feat, target, mask, bf16=False)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-5 * max(1e-3, abs(fd)) + 1e-9, (i, fd, g[i])


def test_transposed_convolution_is_adjoint():
    """<conv(x), d> == <x, conv_transpose(d)> for random x, d at every
    dilation (the data gradient is the adjoint of the forward convolution)."""
    rng = np.random.default_rng(1)
    for dil in R.DIL:
        x = rng.normal(size=(1, 11, 13, 32))
        d = rng.normal(size=(1, 11, 13, 32))
        Wt = rng.normal(size=(32, 9 * 32))
        lhs = np.sum(R.conv3x3(x, Wt, None, dil) * d)
        rhs = np.sum(x * R.conv3x3_transpose(d, Wt, 32, dil))
        assert abs(lhs - rhs) < 1e-9 * max(1.0, abs(lhs)), dil


def test_dilated_convolution_against_direct_sum():
    """conv3x3 at dilation d equals the textbook direct sum
    out[y,x,o] = sum_{dy,dx,c} W[o, (dy+1)*3+dx+1, c] * x[y+d*dy, x+d*dx, c]
    (zero outside), on a small image."""
    rng = np.random.default_rng(2)
    x = rng.normal(size=(1, 9, 10, 3))
    Wt = rng.normal(size=(2, 9 * 3))
    for dil in (1, 2, 4):
        got = R.conv3x3(x, Wt, None, dil)[0]
        want = np.zeros((9, 10, 2))
        for yy in range(9):
            for xx in range(10):
                for dy in (-1, 0, 1):
                    for dx in (-1, 0, 1):
                        sy, sx = yy + dil * dy, xx + dil * dx
                        if 0 <= sy < 9 and 0 <= sx < 10:
                            t = (dy + 1) * 3 + dx + 1
                            want[yy, xx] += Wt[:, t * 3:(t + 1) * 3] @ x[0, sy, sx]
        assert np.allclose(got, want, atol=1e-12), dil


def test_block_inputs_and_targets_are_covered_block_means():
    """Down pass and training targets against float64: the mean of each
    block's covered pixels (7 features, AO), the covered fraction of the
    block's pixels inside the frame (partial blocks at the right and top
    edges), zero inputs and target 1 for blocks with nothing covered."""
    rng = np.random.default_rng(31)
    H, W, f = 37, 50, 4                       # partial blocks on both edges
    feat = rng.normal(0, 1, (H, W, 7)).astype(np.float32)
    cov = rng.random((H, W)) < 0.7
    cov[0:4, 0:4] = False                     # one empty block
    feat[~cov] = 0
    ao = rng.uniform(0, 1, (H, W)).astype(np.float32)
    x, lcov = R.block_inputs(feat, cov, f)
    t, tcov = R.block_targets(ao, cov, f)
    lH, lW = x.shape[:2]
    assert (lH, lW) == (10, 13) and np.array_equal(lcov, tcov)
    for Y in range(lH):
        for X in range(lW):
            blk = (slice(Y * f, min(Y * f + f, H)), slice(X * f, min(X * f + f, W)))
            c = cov[blk]
            npx = c.size
            if not c.any():
                assert not lcov[Y, X] and not x[Y, X].any() and t[Y, X] == 1.0
                continue
            m = feat[blk][c].astype(np.float64).mean(0)
            assert np.allclose(x[Y, X, :7], m, rtol=1e-6, atol=1e-6)
            assert x[Y, X, 7] == np.float32(c.sum()) / np.float32(npx)
            assert abs(t[Y, X] - ao[blk][c].astype(np.float64).mean()) < 1e-6
    assert not x[..., 8:].any()


def test_upsample_is_bilinear_between_block_centres():
    """Up pass: a linear field a + b X + c Y on the network grid comes back
    as a + b u + c w at frame pixel (x, y), u = (x + 0.5)/f - 0.5, wherever
    both neighbours exist (a wrong half-pixel offset or axis swap fails);
    at the edges the nearest block's value; uncovered pixels 1."""
    f, lW, lH = 4, 9, 6
    W, H = lW * f, lH * f
    Y, X = np.mgrid[0:lH, 0:lW]
    a, b, c = 0.3, 0.05, -0.03
    lao = (a + b * X + c * Y).astype(np.float32)
    cov = np.ones((H, W), bool)
    cov[5, 7] = False
    up = R.upsample(lao, cov, f)
    assert up[5, 7] == 1.0
    for y in range(H):
        for x in range(W):
            if not cov[y, x]:
                continue
            u, w = (x + 0.5) / f - 0.5, (y + 0.5) / f - 0.5
            uc, wc = min(max(u, 0), lW - 1), min(max(w, 0), lH - 1)
            assert abs(up[y, x] - (a + b * uc + c * wc)) < 2e-6, (x, y)
    const = R.upsample(np.full((lH, lW), 0.625, np.float32), cov, f)
    assert np.all(const[cov] == np.float32(0.625))


def test_crop_loss_mask_keeps_pixels_the_crop_sees_whole():
    """Crop pixels within the network's reach (sum of the dilations) of a
    crop edge inside the grid are left out of the loss; edges on the grid's
    own border keep their pixels (the whole frame pads there too)."""
    assert R.REACH == 21
    cov = np.ones((64, 64), bool)
    m = R.crop_loss_mask(cov, 10, 10, 64, 200, 100)          # every edge interior
    assert m.sum() == (64 - 42) ** 2 and m[21, 21] and not m[20, 30] and not m[30, 43]
    m = R.crop_loss_mask(cov, 0, 36, 64, 200, 100)           # left and top edges on the grid border
    assert m[:, 0].any() and m[63, 30] and not m[0, 30] and not m[30, 63]
    cov[40, 40] = False
    assert not R.crop_loss_mask(cov, 0, 36, 64, 200, 100)[40, 40]


def test_block_inputs_gbuf_pins():
    """The down pass from the G-buffer (block means of normal and position,
    then the view transform): with V = identity the channels are exactly the
    generic covered-block means of N and of P / 10 (and -Pz / 10); with a
    rigid view transform they equal the block means of the per-pixel
    features in float64 to float32 rounding (the transform is affine); a
    non-finite normal or position, or depth >= 3e38, uncovers the pixel."""
    rng = np.random.default_rng(5)
    H, W, f = 21, 30, 4
    g = rng.normal(0, 1, (19, H, W)).astype(np.float32)
    depth = rng.uniform(0, 1, (H, W)).astype(np.float32)
    depth[rng.random((H, W)) < 0.2] = np.float32(3.4e38)
    depth[3, 5] = depth[10, 11] = np.float32(0.5)
    g[0, 3, 5] = np.nan
    g[4, 10, 11] = np.inf
    cov = R.gbuffer_coverage(depth, g)
    assert not cov[3, 5] and not cov[10, 11] and not cov[depth > 3e38].any()
    assert cov.sum() == ((depth < 3e38).sum() - 2)
    f32 = np.float32
    x, lcov = R.block_inputs_gbuf(depth, g, np.eye(4, dtype=f32), f)
    raw = np.stack([np.where(cov, g[k], 0) for k in range(6)], -1).astype(f32)
    mN, lc2 = R.block_inputs(np.concatenate([raw[..., :3], np.zeros((H, W, 4), f32)], -1), cov, f)
    mP, _ = R.block_inputs(np.concatenate([raw[..., 3:6], np.zeros((H, W, 4), f32)], -1), cov, f)
    assert np.array_equal(lcov, lc2)
    assert np.array_equal(x[..., 1:4], mN[..., 0:3])
    assert np.array_equal(x[..., 4:7], np.where(lcov[..., None], mP[..., 0:3] * f32(0.1), 0))
    assert np.array_equal(x[..., 0], np.where(lcov, -mP[..., 2] * f32(0.1), 0))
    assert np.array_equal(x[..., 7], mN[..., 7])
    # a rigid view transform: rotation about y and x, translation
    a, b = 0.7, -0.3
    Ry = np.array([[np.cos(a), 0, np.sin(a)], [0, 1, 0], [-np.sin(a), 0, np.cos(a)]])
    Rx = np.array([[1, 0, 0], [0, np.cos(b), -np.sin(b)], [0, np.sin(b), np.cos(b)]])
    V = np.eye(4)
    V[:3, :3] = Rx @ Ry
    V[:3, 3] = [0.2, -1.5, -3.0]
    x, lcov = R.block_inputs_gbuf(depth, g, V.astype(f32), f)
    V32 = V.astype(f32).astype(np.float64)
    P = np.concatenate([raw[..., 3:6].astype(np.float64), np.ones((H, W, 1))], -1)
    vp = P @ V32.T
    vn = raw[..., :3].astype(np.float64) @ V32[:3, :3].T
    per = np.concatenate([-vp[..., 2:3] * 0.1, vn, vp[..., :3] * 0.1], -1)
    for Y in range(x.shape[0]):
        for X in range(x.shape[1]):
            blk = (slice(Y * f, min(Y * f + f, H)), slice(X * f, min(X * f + f, W)))
            c = cov[blk]
            if not c.any():
                assert not lcov[Y, X] and not x[Y, X].any()
                continue
            want = per[blk][c].mean(0)
            assert np.allclose(x[Y, X, :7], want, rtol=2e-6, atol=2e-6), (Y, X)
