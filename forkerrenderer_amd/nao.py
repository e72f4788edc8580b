"""Neural ambient occlusion: dataset rendering, training, evaluation and
inference driven through the C ABI (fgl_nao_*).

The network runs on a grid SCALE (4) times coarser than the frame: each
network pixel takes the mean features of its 4 x 4 block of frame pixels,
and the network's AO is interpolated back to the frame pixels.  The
training pairs are G-buffer/AO frames the plain pipeline renders on the
GPU: random layouts of displaced icospheres on a ground plane (the 1280x800
deferred-PBR generator) seen from random camera poses; the target is the
plain kernel SSAO + sigma=2 Gaussian blur, averaged over each block's
covered pixels.  A training step takes 16 random crops of 192 x 192 network
pixels (768 x 768 frame pixels; 4 frames x 4 crops) -- the loss leaves out
the pixels within the network's reach of a crop edge inside the frame --
runs the network's forward and backward kernels and one Adam step (lr
1e-3, betas 0.9/0.999, eps 1e-8).
"""
import ctypes as C
import math
import os

import numpy as np

from . import (PASS_GEOMETRY, SsaoCfg, Context, FglError, lib, make_lookat, make_model_matrix,
               make_perspective, scenegen)


def _bind():
    L = lib()
    if getattr(L, "_nao_bound", False):
        return L
    P = C.c_void_p
    sig = {
        "fgl_nao_param_count": (C.c_int, []),
        "fgl_nao_create": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.POINTER(P)]),
        "fgl_nao_destroy": (None, [P]),
        "fgl_nao_set_state": (C.c_int, [P, P, C.c_int, C.c_int]),
        "fgl_nao_get_state": (C.c_int, [P, P, C.POINTER(C.c_int)]),
        "fgl_nao_stage_crops": (C.c_int, [P, C.c_int, C.c_int, P]),
        "fgl_nao_train_step": (C.c_int, [P, C.c_float, P]),
        "fgl_nao_gradients": (C.c_int, [P, P, P]),
        "fgl_nao_read_batch": (C.c_int, [P, P, P, P]),
        "fgl_nao_infer": (C.c_int, [P]),
        "fgl_nao_set_fused": (C.c_int, [P, C.c_int]),
    }
    for k, (r, a) in sig.items():
        f = getattr(L, k)
        f.restype = r
        f.argtypes = a
    L._nao_bound = True
    return L


def param_count():
    return _bind().fgl_nao_param_count()


SCALE = 4   # frame pixels per network pixel, each way
NIN = 8     # real network inputs: 7 block-mean features + the covered fraction


def init_params(seed):
    """He-normal convolution weights (layer 1 fan-in counts its 8 real input
    channels; the 8 padding channels stay zero), zero biases, head N(0, 1/32),
    head bias 2 so the untrained AO starts near 0.88."""
    rng = np.random.default_rng(seed)
    ch = 32
    w1 = np.zeros((ch, 9, 16))
    w1[:, :, :NIN] = rng.normal(0, math.sqrt(2.0 / (9 * NIN)), (ch, 9, NIN))
    parts = [w1.reshape(-1), np.zeros(ch)]
    for _ in range(3):
        parts += [rng.normal(0, math.sqrt(2.0 / 288), ch * 288), np.zeros(ch)]
    parts += [rng.normal(0, math.sqrt(1.0 / ch), ch), np.array([2.0])]
    p = np.concatenate(parts).astype(np.float32)
    assert p.size == param_count()
    return p


class NeuralAO:
    """The network on the device.  batch x crop x crop network pixels per
    training step; scale: frame pixels per network pixel, each way."""

    def __init__(self, ctx, batch=16, crop=192, seed=0, params=None, scale=SCALE):
        L = _bind()
        self.ctx = ctx
        self.batch, self.crop, self.scale = batch, crop, scale
        self._h = C.c_void_p()
        rc = L.fgl_nao_create(ctx._h, batch, crop, scale, C.byref(self._h))
        if rc != 0:
            raise FglError(f"fgl_nao_create failed: {rc}")
        self.set_params(init_params(seed) if params is None else params)

    def close(self):
        if self._h:
            _bind().fgl_nao_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, p):
        n = param_count()
        st = np.zeros(3 * n, np.float32)
        st[:n] = p
        self._chk(_bind().fgl_nao_set_state(self._h, st.ctypes.data_as(C.c_void_p), 0, 0), "set_state")

    def state(self):
        n = param_count()
        st = np.zeros(3 * n, np.float32)
        step = C.c_int()
        self._chk(_bind().fgl_nao_get_state(self._h, st.ctypes.data_as(C.c_void_p), C.byref(step)), "get_state")
        return st, step.value

    def load_state(self, st, step):
        st = np.ascontiguousarray(st, np.float32)
        self._chk(_bind().fgl_nao_set_state(self._h, st.ctypes.data_as(C.c_void_p), 1, int(step)), "set_state")

    def params(self):
        return self.state()[0][:param_count()].copy()

    def _chk(self, rc, what):
        if rc != 0:
            raise FglError(f"fgl_nao_{what} failed ({rc}): {lib().fgl_last_error(self.ctx._h).decode()}")

    def stage(self, slot0, xy):
        xy = np.ascontiguousarray(np.asarray(xy, np.int32).reshape(-1, 2))
        self._chk(_bind().fgl_nao_stage_crops(self._h, slot0, len(xy), xy.ctypes.data_as(C.c_void_p)), "stage")

    def train_step(self, lr=1e-3):
        st = np.zeros(3, np.float32)
        self._chk(_bind().fgl_nao_train_step(self._h, lr, st.ctypes.data_as(C.c_void_p)), "train_step")
        return st  # sum sq err, sum abs err, covered pixels

    def gradients(self):
        g = np.zeros(param_count(), np.float32)
        st = np.zeros(3, np.float32)
        self._chk(_bind().fgl_nao_gradients(self._h, g.ctypes.data_as(C.c_void_p), st.ctypes.data_as(C.c_void_p)),
                  "gradients")
        return g, st

    def read_batch(self):
        px = self.batch * self.crop * self.crop
        feat = np.zeros(px * 16, np.uint16)
        tgt = np.zeros(px, np.float32)
        mask = np.zeros(px, np.uint8)
        self._chk(_bind().fgl_nao_read_batch(self._h, feat.ctypes.data_as(C.c_void_p), tgt.ctypes.data_as(C.c_void_p),
                                             mask.ctypes.data_as(C.c_void_p)), "read_batch")
        s = (self.batch, self.crop, self.crop)
        return feat.reshape(*s, 16), tgt.reshape(s), mask.reshape(s)

    def infer(self):
        """Neural AO of the context's current frame into its AO target."""
        self._chk(_bind().fgl_nao_infer(self._h), "infer")

    def set_fused(self, fused):
        """Inference through the two fused kernels (True, the default) or one
        kernel per layer (False); the results are the same bits."""
        self._chk(_bind().fgl_nao_set_fused(self._h, int(bool(fused))), "set_fused")


# ---------------- training data: the plain pipeline on random scenes ----------------

def ssao_cfg(seed=scenegen.SSAO_KERNEL_SEED, n=64, radius=0.5, sigma=2.0):
    """Config-2 SSAO: 64-sample kernel r=0.5, bias 5e-4, range 1e-2, 4x4 noise,
    sigma=2 separable blur (the same sets the .scene keys produce)."""
    import tempfile
    from . import Scene
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "cfg.scene")
        with open(p, "w") as fh:
            fh.write(f"light point 0 5 0 1 1 1\nssaokernel {n} {radius} 0.0005 0.01 {seed}\n"
                     f"aoblur gaussian {sigma}\n")
        sc = Scene(p)
        return SsaoCfg.from_buffer_copy(sc.info.ssao)


class SceneSet:
    """Assets of the deferred-PBR scene generator (displaced icosphere of
    radius 1, ground plane, 2048^2 albedo / roughness / metalness / normal
    maps with generated tangents), uploaded once into a context.  The normal
    map's size matters: minified without mipmaps, it sets how rough the
    per-pixel normals are, and with them the SSAO the network learns."""

    def __init__(self, ctx, workdir, subdiv=5, seed=100, tex_size=2048):
        from . import Scene
        os.makedirs(workdir, exist_ok=True)
        scenegen._pbr_scene_assets(workdir, seed, subdiv, tex_size)
        p = os.path.join(workdir, "assets.scene")
        with open(p, "w") as fh:
            fh.write(f"mode deferred\nlight point 2 5 5 3 3 3\ntexture repeat linear\n"
                     f"model {workdir}/ground.obj false true 0 0 0 0 1\n"
                     f"model {workdir}/sphere.obj false true 0 0 0 0 1\n")
        (self.ground, _), (self.sphere, _) = ctx.upload_scene(Scene(p))


def layout(seed):
    """A random scene, seeded: a ground plane of half-size 4 and displaced
    icospheres in one of four arrangements -- one large sphere (radius
    0.8-1.2), 1-6 scattered spheres (radius 0.3-1.0), an n x n grid (n 3-8,
    spacing 0.8) of small spheres (radius 0.15-0.45) resting on or sunk into
    the ground, or the heavy benchmark scene's family: an n x n grid (n 6-9,
    spacing 0.8) of spheres of radius 0.3 x [0.5, 1.5] with their centres
    0.45 above the ground."""
    rng = np.random.default_rng(seed)
    kind = int(rng.integers(0, 4))
    objs = []
    if kind == 0:
        r = float(rng.uniform(0.8, 1.2))
        objs.append(((0.0, float(rng.uniform(0.55, 1.0)) * r, 0.0), float(rng.uniform(0, 360)), r))
    elif kind == 1:
        for _ in range(int(rng.integers(1, 7))):
            r = float(rng.uniform(0.3, 1.0))
            x, z = rng.uniform(-2.2, 2.2, 2)
            objs.append(((float(x), float(rng.uniform(0.55, 1.0)) * r, float(z)), float(rng.uniform(0, 360)), r))
    elif kind == 2:
        n = int(rng.integers(3, 9))
        for gz in range(n):
            for gx in range(n):
                r = float(rng.uniform(0.15, 0.45))
                x, z = (gx - (n - 1) / 2) * 0.8, (gz - (n - 1) / 2) * 0.8
                objs.append(((x, float(rng.uniform(0.55, 1.6)) * r, z), float(rng.uniform(0, 360)), r))
    else:
        n = int(rng.integers(6, 10))
        for gz in range(n):
            for gx in range(n):
                r = 0.3 * float(rng.uniform(0.5, 1.5))
                x, z = (gx - (n - 1) / 2) * 0.8, (gz - (n - 1) / 2) * 0.8
                objs.append(((x, 0.45, z), float(rng.uniform(0, 360)), r))
    ground = make_model_matrix((0.0, 0.0, 0.0), 0.0, 4.0)
    return ground, [make_model_matrix(t, rot, r) for t, rot, r in objs]


def camera(seed):
    """Random orbit pose: distance 3.5-7, elevation 12-55 degrees, any azimuth,
    target jittered around the origin."""
    rng = np.random.default_rng(seed)
    r = rng.uniform(3.5, 7.0)
    el = math.radians(rng.uniform(12, 55))
    az = rng.uniform(0, 2 * math.pi)
    eye = (r * math.cos(el) * math.cos(az), r * math.sin(el), r * math.cos(el) * math.sin(az))
    at = tuple(rng.uniform(-0.6, 0.6, 3) * (1, 0.3, 1))
    return eye, at


def render_frame(ctx, sset, frame_seed, W, H, cfg, plain_ao=True):
    """G-buffer pass of a random layout + camera; optionally the plain SSAO +
    blur into the AO target.  Shadows off (the AO pass does not use them)."""
    ground, objs = layout(frame_seed // 1000)
    eye, at = camera(frame_seed)
    V = make_lookat(eye, at)
    P = make_perspective(45.0, np.float32(W) / np.float32(H), 0.01, 20.0)
    from . import MAT_PROJ, MAT_VIEW
    ctx.set_matrix(MAT_VIEW, V)
    ctx.set_matrix(MAT_PROJ, P)
    ctx.set_eye(eye)
    ctx.set_shadow(False)
    ctx.set_viewport(0, 0, W, H)
    ctx.init_targets(W, H)
    ctx.begin_pass(PASS_GEOMETRY)
    ctx.draw_model(sset.ground, ground)
    for M in objs:
        ctx.draw_model(sset.sphere, M)
    ctx.end_pass()
    if plain_ao:
        ctx.ssao(cfg)


def frame_size(frame_seed):
    """Training frames: one pose in four of every layout is rendered at
    1920 x 1080, the others at 1280 x 800."""
    return (1920, 1080) if (frame_seed % 1000) % 4 == 3 else (1280, 800)


def frame_seeds(split, n):
    """Frame seeds layout * 1000 + pose.  Training: layouts 0..399, frame i
    is layout i % 400, pose i // 400 (4000 frames = 10 poses each).
    Evaluation: held-out layouts 1000..1019, frame i is layout 1000 + i % 20,
    pose i // 20 (200 frames = 10 poses each)."""
    if split == "train":
        return [(i % 400) * 1000 + i // 400 for i in range(n)]
    return [(1000 + i % 20) * 1000 + i // 20 for i in range(n)]


def layout_kind(layout_seed):
    """The arrangement `layout` draws for a layout seed (0: one sphere,
    1: scattered, 2: small-sphere grid, 3: the benchmark scene's grid)."""
    return int(np.random.default_rng(layout_seed).integers(0, 4))


def train_frames(frames, favour_grid=1):
    """The training frame seeds, the grid layouts (kinds 2 and 3) listed
    `favour_grid` times."""
    fl = frame_seeds("train", frames)
    if favour_grid > 1:
        fl = fl + [f for f in fl if layout_kind(f // 1000) >= 2] * (favour_grid - 1)
    return fl


def step_batch(step, seed, frames, crop, batch, scale=SCALE, favour_grid=1):
    """The frames and crop origins of training step `step`: a pure function
    of (seed, step), so a run resumed from a checkpoint sees the same data.
    4 frames per step, batch/4 crops each; crop origins in network pixels of
    each frame's network grid (frame_size / scale)."""
    order = np.random.default_rng(seed).permutation(train_frames(frames, favour_grid))
    rng = np.random.default_rng([seed, step])
    per = batch // 4
    out = []
    for j in range(4):
        fs = int(order[(4 * step + j) % len(order)])
        W, H = frame_size(fs)
        lW, lH = -(-W // scale), -(-H // scale)
        xy = np.stack([rng.integers(0, lW - crop + 1, per), rng.integers(0, lH - crop + 1, per)], 1)
        out.append((fs, xy))
    return out


def lr_at(step, total, base=1e-3):
    """Adam learning rate: `base` for the first 3/4 of the run, then a linear
    decay to base/100 at the last step."""
    t0 = 0.75 * total
    if total <= 0 or step < t0:
        return base
    return base * (1.0 - 0.99 * min(1.0, (step - t0) / max(total - t0, 1.0)))


def train(nao, ctx, sset, steps, start=0, total=None, seed=0, frames=4000, lr=1e-3, log=None, cfg=None,
          favour_grid=1):
    """Steps start .. start+steps-1 of a training run of `total` steps
    (default start+steps); returns per-step (sum sq err, sum |err|, covered
    network pixels).  favour_grid: how many times the grid layouts are
    listed among the training frames."""
    cfg = cfg if cfg is not None else ssao_cfg()
    total = start + steps if total is None else total
    per = nao.batch // 4
    hist = []
    for step in range(start, start + steps):
        for j, (fs, xy) in enumerate(step_batch(step, seed, frames, nao.crop, nao.batch, nao.scale, favour_grid)):
            W, H = frame_size(fs)
            render_frame(ctx, sset, fs, W, H, cfg)
            nao.stage(j * per, xy)
        sse, sae, cnt = nao.train_step(lr_at(step, total, lr))
        hist.append((float(sse), float(sae), float(cnt)))
        if log and (step % 500 == 0 or step == start + steps - 1):
            k = hist[-min(len(hist), 100):]
            n = max(sum(h[2] for h in k), 1)
            log(f"step {step}: lr {lr_at(step, total, lr):.2e} mse(100) {sum(h[0] for h in k) / n:.5f} "
                f"mae(100) {sum(h[1] for h in k) / n:.4f}")
    return hist


def save_checkpoint(path, nao, seed, frames, W=1280, H=800, evaluation=None):
    """Weights, Adam moments, step and the data recipe (seed, frame count,
    frame size, crop, batch) in one .npz, plus the held-out evaluation as
    JSON text when given."""
    import json
    st, step = nao.state()
    np.savez(path, state=st, step=np.int64(step), seed=np.int64(seed), frames=np.int64(frames),
             size=np.array([W, H], np.int64), crop=np.int64(nao.crop), batch=np.int64(nao.batch),
             scale=np.int64(nao.scale), nparam=np.int64(param_count()),
             eval_json=np.array(json.dumps(evaluation or {})))


def load_checkpoint(path):
    z = np.load(path)
    if int(z["nparam"]) != param_count():
        raise FglError(f"{path}: parameter count {int(z['nparam'])} != {param_count()}")
    if "scale" not in z.files:
        raise FglError(f"{path}: a checkpoint of the full-resolution network (no network scale recorded)")
    return {k: z[k] for k in z.files}


DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "nao_weights.npz")


def load_weights(path=DEFAULT_WEIGHTS):
    """Trained parameters (flat float32) from a checkpoint."""
    return load_checkpoint(path)["state"][:param_count()].copy()


def audit_fallback(nao_ao, plain_ao, cover, frame_index, tile=64, every=16, threshold=0.05):
    """1-in-`every` tile audit: tiles t with t % every == frame_index % every
    are compared with plain SSAO; a tile whose MAE over covered pixels exceeds
    `threshold` falls back to the plain result.  Returns (ao, audited, failed)."""
    H, W = nao_ao.shape
    out = nao_ao.copy()
    audited = failed = 0
    ty, tx = (H + tile - 1) // tile, (W + tile - 1) // tile
    for t in range(ty * tx):
        if t % every != frame_index % every:
            continue
        y0, x0 = (t // tx) * tile, (t % tx) * tile
        sl = (slice(y0, y0 + tile), slice(x0, x0 + tile))
        c = cover[sl]
        if not c.any():
            continue
        audited += 1
        if np.abs(nao_ao[sl] - plain_ao[sl])[c].mean() > threshold:
            failed += 1
            out[sl] = plain_ao[sl]
    return out, audited, failed


def evaluate(nao, ctx, sset, n_frames=200, W=1280, H=800, cfg=None, unblurred=False):
    """MAE / max error of neural AO against plain SSAO+blur on held-out frames
    (covered pixels), the audit fallback rate and the MAE after fallback.
    With unblurred=True also the MAE of SSAO without its blur against the
    same target (the cheaper plain alternative)."""
    from . import TARGET_AO, TARGET_DEPTH
    cfg = cfg if cfg is not None else ssao_cfg()
    raw = None
    if unblurred:
        raw = SsaoCfg.from_buffer_copy(cfg)
        raw.blur = 0
    sae = sae_fb = sae_raw = cnt = 0.0
    mx = 0.0
    audited = failed = 0
    for i, fs in enumerate(frame_seeds("eval", n_frames)):
        if raw is not None:
            render_frame(ctx, sset, fs, W, H, raw)
            noblur = ctx.read(TARGET_AO).reshape(H, W)
            ctx.ssao(cfg)
        else:
            render_frame(ctx, sset, fs, W, H, cfg)
        plain = ctx.read(TARGET_AO).reshape(H, W)
        cover = ctx.read(TARGET_DEPTH).reshape(H, W) < 3e38
        nao.infer()
        pred = ctx.read(TARGET_AO).reshape(H, W)
        err = np.abs(pred - plain)[cover]
        sae += float(err.sum())
        cnt += err.size
        mx = max(mx, float(err.max()) if err.size else 0.0)
        fixed, a, f = audit_fallback(pred, plain, cover, i)
        sae_fb += float(np.abs(fixed - plain)[cover].sum())
        if raw is not None:
            sae_raw += float(np.abs(noblur - plain)[cover].sum())
        audited += a
        failed += f
    out = {"mae": sae / max(cnt, 1), "max_err": mx, "fallback_rate": failed / max(audited, 1),
           "mae_after_fallback": sae_fb / max(cnt, 1), "frames": n_frames, "audited_tiles": audited}
    if raw is not None:
        out["mae_unblurred_ssao"] = sae_raw / max(cnt, 1)
    return out


def main(argv=None):
    """python -m forkerrenderer_amd.nao --steps N --out ckpt.npz: trains from
    scratch (or --resume), evaluates on the held-out frames, writes the
    checkpoint and a JSON line of metrics."""
    import argparse
    import json
    import time
    ap = argparse.ArgumentParser()
    ap.add_argument("--total", type=int, default=8000, help="steps of the whole run (the lr schedule's length)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--frames", type=int, default=4000)
    ap.add_argument("--eval-frames", type=int, default=200)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--out", required=True)
    ap.add_argument("--resume", default=None)
    ap.add_argument("--workdir", default="_scenes/_nao")
    ap.add_argument("--favour-grid", type=int, default=1,
                    help="list the grid layouts this many times among the training frames")
    a = ap.parse_args(argv)
    ctx = Context(0)
    sset = SceneSet(ctx, a.workdir)
    cfg = ssao_cfg()
    nao = NeuralAO(ctx, seed=a.seed)
    start = 0
    if a.resume:
        ck = load_checkpoint(a.resume)
        nao.load_state(ck["state"], int(ck["step"]))
        start = int(ck["step"])
    base = evaluate(nao, ctx, sset, min(a.eval_frames, 40), cfg=cfg) if start == 0 else None
    t0 = time.time()
    hist = []
    chunk = 2000
    for s0 in range(start, a.total, chunk):
        n = min(chunk, a.total - s0)
        hist += train(nao, ctx, sset, n, start=s0, total=a.total, seed=a.seed, frames=a.frames, lr=a.lr, cfg=cfg,
                      log=lambda s: print(s, f"({time.time() - t0:.1f}s)", flush=True), favour_grid=a.favour_grid)
        ev = evaluate(nao, ctx, sset, 20, cfg=cfg)
        print(f"step {s0 + n}: held-out MAE (20 frames) {ev['mae']:.4f}", flush=True)
    t_train = time.time() - t0
    ev = evaluate(nao, ctx, sset, a.eval_frames, cfg=cfg, unblurred=True)
    save_checkpoint(a.out, nao, a.seed, a.frames, evaluation=ev)
    rec = {"steps": a.total, "train_s": t_train, "final_train_mse": hist[-1][0] / max(hist[-1][2], 1),
           "eval": ev, "untrained_eval40": base}
    print(json.dumps(rec), flush=True)
    nao.close()
    ctx.close()


if __name__ == "__main__":
    main()
