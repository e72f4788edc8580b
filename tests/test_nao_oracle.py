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
