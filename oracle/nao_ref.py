"""CPU oracle of the neural-AO network (TEST INFRASTRUCTURE).

The network runs on a grid `f` times coarser than the frame;
block_inputs_gbuf, block_targets and upsample restate the device's down
pass, training targets and up pass in its float32 operation order
(block_inputs is the generic covered-block mean).

Forward and backward of: conv3x3 16->32 ReLU, 3x conv3x3 32->32 ReLU,
conv1x1 32->1 sigmoid, 'same' zero padding, the four 3x3 convolutions
dilated by DIL = (1, 2, 6, 12) (taps at offsets -d, 0, +d: receptive field
43 x 43), loss = sum over covered pixels of (y - t)^2 / n_pixels.  Sums are evaluated in float64.  With bf16=True the
values the device keeps in BF16 (weights of layers 1-4, activations, data
gradients) are rounded to BF16 at the same points; with bf16=False the
network is plain float64, which the finite-difference test uses to pin the
backward formulas.

Parameter layout (flat, float32 on the device): W1 [32][9][16], b1[32],
W2..W4 [32][9][32], b2..b4 [32], W5 [32], b5 [1]; k = tap*Cin + ci with
tap = (dy+1)*3 + (dx+1).
"""
import numpy as np

C0, CH = 16, 32
K1, K2 = 9 * C0, 9 * CH
SIZES = [("W1", CH * K1), ("b1", CH), ("W2", CH * K2), ("b2", CH), ("W3", CH * K2), ("b3", CH),
         ("W4", CH * K2), ("b4", CH), ("W5", CH), ("b5", 1)]
NPARAM = sum(s for _, s in SIZES)
DIL = (1, 2, 6, 12)


def split(params):
    out, o = {}, 0
    for name, s in SIZES:
        out[name] = np.asarray(params[o:o + s], np.float64)
        o += s
    return out


def bf16_round(x):
    """float -> bf16 (round to nearest even) -> float, as the device's f2bf."""
    x32 = np.ascontiguousarray(np.asarray(x, np.float32))
    u = x32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFFFFFF
    r = ((u >> 16) << 16).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def bf16_bits_to_float(b):
    return (np.asarray(b, np.uint32) << 16).view(np.float32).astype(np.float64)


def _patches(x, dil=1):
    """x: [N,H,W,C] -> [N,H,W,9,C] with zero padding (tap order dy, dx; tap
    offsets -dil, 0, +dil)."""
    N, H, W, Cc = x.shape
    d = dil
    xp = np.zeros((N, H + 2 * d, W + 2 * d, Cc), x.dtype)
    xp[:, d:d + H, d:d + W] = x
    taps = [xp[:, d + dy * d:d + dy * d + H, d + dx * d:d + dx * d + W] for dy in (-1, 0, 1) for dx in (-1, 0, 1)]
    return np.stack(taps, axis=3)


def conv3x3(x, W, b, dil=1):
    """x [N,H,W,Cin], W [Co][9*Cin] -> [N,H,W,Co]"""
    Co = W.shape[0]
    p = _patches(x, dil).reshape(x.shape[0], x.shape[1], x.shape[2], -1)
    out = p @ W.reshape(Co, -1).T
    return out + (b if b is not None else 0.0)


def conv3x3_transpose(d, W, cin, dil=1):
    """Gradient w.r.t. the input of conv3x3: d [N,H,W,Co], W [Co][9*cin]."""
    Co = W.shape[0]
    Wr = W.reshape(Co, 9, cin)
    Wf = Wr[:, ::-1, :].transpose(2, 1, 0).reshape(cin, 9 * Co)   # flipped taps, transposed
    return conv3x3(d, Wf, None, dil)


def forward(params, feat, bf16=True):
    P = split(params)
    r = bf16_round if bf16 else (lambda a: np.asarray(a, np.float64))
    x = np.asarray(feat, np.float64)
    acts = []
    for l in range(4):
        W = r(P[f"W{l + 1}"]).reshape(CH, -1)
        z = conv3x3(x, W, P[f"b{l + 1}"], DIL[l])
        a = r(np.maximum(z, 0.0))
        acts.append(a)
        x = a
    z5 = acts[-1] @ P["W5"] + P["b5"][0]
    y = 1.0 / (1.0 + np.exp(-z5))
    return acts, y


def loss_and_grads(params, feat, target, mask, bf16=True):
    """Returns (loss, flat gradient float64, stats (sse, sae, count))."""
    P = split(params)
    r = bf16_round if bf16 else (lambda a: np.asarray(a, np.float64))
    acts, y = forward(params, feat, bf16)
    n = y.size
    m = np.asarray(mask, bool)
    t = np.asarray(target, np.float64)
    diff = y - t
    loss = float(np.sum(np.where(m, diff * diff, 0.0)) / n)
    dz = np.where(m, 2.0 * diff / n, 0.0) * y * (1.0 - y)
    g = {}
    a4 = acts[3]
    g["W5"] = np.einsum("nhw,nhwc->c", dz, a4)
    g["b5"] = np.array([dz.sum()])
    d = r(np.where(a4 > 0, dz[..., None] * P["W5"], 0.0))
    ins = [np.asarray(feat, np.float64)] + acts[:3]
    for l in range(3, -1, -1):
        cin = C0 if l == 0 else CH
        pat = _patches(ins[l], DIL[l]).reshape(d.shape[0], d.shape[1], d.shape[2], -1)
        g[f"W{l + 1}"] = np.einsum("nhwo,nhwk->ok", d, pat).reshape(-1)
        g[f"b{l + 1}"] = d.sum(axis=(0, 1, 2))
        if l > 0:
            W = r(P[f"W{l + 1}"]).reshape(CH, -1)
            dx = conv3x3_transpose(d, W, cin, DIL[l])
            d = r(np.where(ins[l] > 0, dx, 0.0))
    flat = np.concatenate([np.asarray(g[name], np.float64).reshape(-1) for name, _ in SIZES])
    stats = (float(np.sum(np.where(m, diff * diff, 0))), float(np.sum(np.where(m, np.abs(diff), 0))),
             float(m.sum()))
    return loss, flat, stats


NIN = 8   # real inputs: 7 block-mean features + the covered fraction


def init_params(seed):
    """He-normal init for the convolutions (fan-in counts the 8 real input
    channels of layer 1), small head, b5 = 2 (AO starts near 0.88)."""
    rng = np.random.default_rng(seed)
    parts = []
    w1 = np.zeros((CH, 9, C0))
    w1[:, :, :NIN] = rng.normal(0, np.sqrt(2.0 / (9 * NIN)), (CH, 9, NIN))
    parts += [w1.reshape(-1), np.zeros(CH)]
    for _ in range(3):
        parts += [rng.normal(0, np.sqrt(2.0 / K2), CH * K2), np.zeros(CH)]
    parts += [rng.normal(0, np.sqrt(1.0 / CH), CH), np.array([2.0])]
    return np.concatenate(parts).astype(np.float32)


# ---------------- down pass, targets, up pass (float32, device order) ----------------

def _blocks(W, H, f):
    return -(-W // f), -(-H // f)


def block_inputs(feat, cov, f):
    """Network inputs from per-pixel features.  feat: (H, W, 7) float32 pixel
    features (zero where uncovered), cov: (H, W) bool.  Per network pixel
    (X, Y): the covered pixels of its block summed in row-major order from
    0, times 1/n (n covered); channel 7 = n / (block pixels in the frame).
    Returns (lH, lW, 16) float32 and the network coverage (n > 0)."""
    H, W = cov.shape
    lW, lH = _blocks(W, H, f)
    f32 = np.float32
    acc = np.zeros((lH, lW, 7), f32)
    n = np.zeros((lH, lW), np.int64)
    m = np.zeros((lH, lW), np.int64)
    for dy in range(f):
        for dx in range(f):
            ys, xs = np.arange(lH) * f + dy, np.arange(lW) * f + dx
            yv, xv = ys < H, xs < W
            sub = np.zeros((lH, lW, 7), f32)
            c = np.zeros((lH, lW), bool)
            yy, xx = ys[yv], xs[xv]
            sub[np.ix_(yv, xv)] = feat[np.ix_(yy, xx)]
            c[np.ix_(yv, xv)] = cov[np.ix_(yy, xx)]
            acc = np.where(c[..., None], acc + sub, acc)
            n += c
            m[np.ix_(yv, xv)] += 1
    inv = np.where(n > 0, f32(1) / np.maximum(n, 1).astype(f32), f32(0))
    out = np.zeros((lH, lW, 16), f32)
    out[..., :7] = acc * inv[..., None]
    out[..., 7] = np.where(m > 0, n.astype(f32) / np.maximum(m, 1).astype(f32), f32(0))
    return out, n > 0


def gbuffer_coverage(depth, gbuf):
    """Network coverage of each frame pixel: something drawn (depth below
    3e38) with a finite G-buffer normal and position (planes 0-5)."""
    with np.errstate(invalid="ignore"):
        cov = np.asarray(depth) < 3e38
        for k in range(6):
            cov = cov & np.isfinite(gbuf[k])
    return cov


def _dot(row, v):
    """Dot product in float32, summed from 0 in order (the device's dot3/dot4)."""
    f32 = np.float32
    r = np.zeros_like(v[0], dtype=f32)
    for a, b in zip(row, v):
        r = r + f32(a) * b
    return r


def block_inputs_gbuf(depth, gbuf, V, f):
    """The device's down pass: per network pixel the block means of the
    covered pixels' G-buffer normal N and position P (planes 0-5; summed in
    row-major order from 0, times 1/n), then the view-space features of the
    means -- depth -z/10, V3 N, V (P, 1) / 10 (V row-major 4x4; one dot
    product per row) -- and the covered fraction.  The view transform is
    affine, so this is block_inputs of the per-pixel features up to
    rounding.  depth (H, W), gbuf (>= 6, H, W) float32.  Returns
    (lH, lW, 16) float32 and the network coverage."""
    f32 = np.float32
    cov = gbuffer_coverage(depth, gbuf)
    raw = np.stack([np.where(cov, gbuf[k], f32(0)).astype(f32) for k in range(6)], -1)
    H, W = cov.shape
    lW, lH = _blocks(W, H, f)
    acc = np.zeros((lH, lW, 6), f32)
    n = np.zeros((lH, lW), np.int64)
    m = np.zeros((lH, lW), np.int64)
    for dy in range(f):
        for dx in range(f):
            ys, xs = np.arange(lH) * f + dy, np.arange(lW) * f + dx
            yv, xv = ys < H, xs < W
            sub = np.zeros((lH, lW, 6), f32)
            c = np.zeros((lH, lW), bool)
            yy, xx = ys[yv], xs[xv]
            sub[np.ix_(yv, xv)] = raw[np.ix_(yy, xx)]
            c[np.ix_(yv, xv)] = cov[np.ix_(yy, xx)]
            acc = np.where(c[..., None], acc + sub, acc)
            n += c
            m[np.ix_(yv, xv)] += 1
    inv = np.where(n > 0, f32(1) / np.maximum(n, 1).astype(f32), f32(0))
    mean = acc * inv[..., None]
    Vm = np.asarray(V, f32).reshape(4, 4)
    N = [mean[..., 0], mean[..., 1], mean[..., 2]]
    P = [mean[..., 3], mean[..., 4], mean[..., 5], np.ones_like(mean[..., 0])]
    vp = [_dot(Vm[r], P) for r in range(3)]
    vn = [_dot(Vm[r, :3], N) for r in range(3)]
    feat = [-vp[2] * f32(0.1), vn[0], vn[1], vn[2], vp[0] * f32(0.1), vp[1] * f32(0.1), vp[2] * f32(0.1)]
    out = np.zeros((lH, lW, 16), f32)
    for k in range(7):
        out[..., k] = np.where(n > 0, feat[k], f32(0))
    out[..., 7] = np.where(m > 0, n.astype(f32) / np.maximum(m, 1).astype(f32), f32(0))
    return out, n > 0


def block_targets(ao, cov, f):
    """Training targets: per network pixel the covered pixels' AO summed in
    row-major order from 0, times 1/n; 1 where no pixel is covered."""
    H, W = cov.shape
    lW, lH = _blocks(W, H, f)
    f32 = np.float32
    acc = np.zeros((lH, lW), f32)
    n = np.zeros((lH, lW), np.int64)
    for dy in range(f):
        for dx in range(f):
            ys, xs = np.arange(lH) * f + dy, np.arange(lW) * f + dx
            yv, xv = ys < H, xs < W
            a = np.zeros((lH, lW), f32)
            c = np.zeros((lH, lW), bool)
            a[np.ix_(yv, xv)] = ao[np.ix_(ys[yv], xs[xv])]
            c[np.ix_(yv, xv)] = cov[np.ix_(ys[yv], xs[xv])]
            acc = np.where(c, acc + a, acc)
            n += c
    inv = np.where(n > 0, f32(1) / np.maximum(n, 1).astype(f32), f32(0))
    return np.where(n > 0, acc * inv, f32(1)), n > 0


def upsample(lao, cov, f):
    """Frame AO from the network grid's AO: at covered pixel (x, y), bilinear
    weights at u = (x + 0.5)/f - 0.5, w = (y + 0.5)/f - 0.5 between the
    clamped neighbours floor(u), floor(u) + 1 (and in y); 1 elsewhere."""
    H, W = cov.shape
    lH, lW = lao.shape
    f32 = np.float32
    inv = f32(1) / f32(f)
    u = (np.arange(W, dtype=f32) + f32(0.5)) * inv - f32(0.5)
    w = (np.arange(H, dtype=f32) + f32(0.5)) * inv - f32(0.5)
    fu, fw = np.floor(u), np.floor(w)
    tx, ty = (u - fu).astype(f32), (w - fw).astype(f32)
    xa, xb = np.clip(fu.astype(int), 0, lW - 1), np.clip(fu.astype(int) + 1, 0, lW - 1)
    ya, yb = np.clip(fw.astype(int), 0, lH - 1), np.clip(fw.astype(int) + 1, 0, lH - 1)
    L = lao.astype(f32)
    top = L[ya][:, xa] * (f32(1) - tx)[None, :] + L[ya][:, xb] * tx[None, :]
    bot = L[yb][:, xa] * (f32(1) - tx)[None, :] + L[yb][:, xb] * tx[None, :]
    out = top * (f32(1) - ty)[:, None] + bot * ty[:, None]
    return np.where(cov, out, f32(1)).astype(f32)


REACH = sum(DIL)   # receptive-field radius in network pixels


def crop_loss_mask(cov, x0, y0, S, lW, lH):
    """Training mask of a crop: covered, and at least REACH pixels inside
    every crop edge that is not an edge of the lW x lH network grid."""
    cy, cx = np.mgrid[0:S, 0:S]
    inside = ((x0 == 0) | (cx >= REACH)) & ((x0 + S == lW) | (cx < S - REACH)) & \
             ((y0 == 0) | (cy >= REACH)) & ((y0 + S == lH) | (cy < S - REACH))
    return cov & inside
