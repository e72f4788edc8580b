"""Benchmark: deferred PBR + SSAO frames/s at 1080p on one MI355X.

Workload (default): config 3 of the benchmark set -- 1920x1080 with SSAA 2x
(3840x2160 internal), 65 models (8x8 displaced icospheres of 81,920 faces +
ground plane = 5,242,882 triangles), 2048^2 shadow map with PCSS (16 blocker
/ 32 PCF samples, area 0.05), 64-sample kernel SSAO + sigma=2 Gaussian blur,
a point and a directional light.  One step = one full frame: shadow pass,
geometry pass, SSAO + blur, lighting, quantise + 2x2 SSAA resolve, with every
input resident in HBM.

Prints one JSON line (see the repository's contract).  The CPU baseline is the
oracle (the C restatement of the reference path) rendering the same whole
frame single-threaded on one pinned host core, after the GPU work.
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "deferred PBR+SSAO frames/s @1080p; neural-AO vs plain ms/frame at AO MAE≤0.02"
HBM_PEAK_GBS = 8000.0


def gen_scene(cfg):
    from forkerrenderer_amd import scenegen as G
    d = os.path.join(ROOT, "_scenes", cfg)
    p = os.path.join(d, f"{cfg}.scene")
    if not os.path.exists(p):
        {"c1": G.config1, "c2": G.config2, "c3": G.config3, "c4": G.config4}[cfg](d)
    return p


def scene_stats(sc):
    I = sc.info
    ntri = 0
    npos = 0
    for i in range(I.n_models):
        m, _ = sc.model(i)
        ntri += len(m["vi"])
        npos += len(m["pos"])
    return ntri, npos


def algorithmic_bytes(name, I, ntri, npos):
    """Compulsory HBM bytes of one launch of each kernel group (FP32 SoA)."""
    k = I.ssaa_k if I.ssaa_on else 1
    P = I.width * k * I.height * k
    S = I.shadow_map_size if I.shadow_map_size > 0 else None
    SP = S * S if S else P
    tab = {
        "raster": 8 * P + 16 * ntri + 16 * npos,                  # depth + winner write; triangle and vertex records
        "gbuffer_resolve": 80 * P + 12 * ntri + 44 * npos,        # winner read, G-buffer write; indices, vertices
        "forward_resolve": 16 * P + 12 * ntri + 44 * npos,        # winner read, colour write; indices, vertices
        "raster_shadow": 8 * SP + 16 * ntri + 16 * npos,          # shadow depth + value; triangle, vertex records
        "lighting": 92 * P + 4 * SP,                              # 80 read + 12 write, shadow map once
        "ssao": 32 * P,                                           # normal, position, depth, AO
        "blur_gaussian": 8 * P,
        "blur_reference": 16 * P,
        "resolve": 12 * P + 3 * (P // (k * k)),
    }
    return tab.get(name)


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC
    summary (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, calibrated on a plain
    copy), or None."""
    import re
    p = os.path.join(ROOT, "profiles", f"{kernel}_c3_hbm_traffic.md")
    if not os.path.exists(p):
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*(\d+)", open(p).read())
    return int(m.group(1)) if m else None


def run_gpu(sc, steps, warmup, tile, tiles=None, dominant=None):
    """Per-kernel breakdown (every kernel timed with HIP events, a separate
    run of the same frames), then the throughput run: `steps` frames with
    events only around the dominant kernel (chosen by `dominant` from the
    breakdown), whose launches the roofline uses."""
    import forkerrenderer_amd as F
    from forkerrenderer_amd.autotune import autotune
    ctx = F.Context(0)
    tuned = None
    if tiles:
        ctx.set_pass_tile_size(F.PASS_SHADOW, tiles[0])
        ctx.set_pass_tile_size(F.PASS_GEOMETRY, tiles[1])
        tuned = ({"shadow": tiles[0], "main": tiles[1]}, None)
    elif tile:
        ctx.set_tile_size(tile)
    else:
        tuned = autotune(ctx, sc)
        print(f"[bench] autotune {tuned[0]} {tuned[1]}", file=sys.stderr, flush=True)
    for _ in range(warmup):
        ctx.render_scene(sc)
    ctx.sync()

    def frames(n, only):
        ctx.read_timings()
        ctx.enable_timing(True, only)
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(n):
            ctx.render_scene(sc)
        ctx.sync()
        t1 = time.perf_counter()
        timings = ctx.read_timings()
        ctx.enable_timing(False)
        per = {}
        for name, ms in timings:
            per.setdefault(name, []).append(ms)
        return (t1 - t0) / n, per

    nb = max(5, steps // 2)
    sec_all, per = frames(nb, None)
    dom = dominant(per, nb)
    sec, per_dom = frames(steps, dom)
    img = ctx.read(F.TARGET_IMAGE)
    ctx.close()
    return sec, per, nb, sec_all, dom, per_dom[dom], img, tuned


def run_neural_ao(steps, warmup, heldout_frames=200):
    """Neural AO against plain SSAO + blur on the 1920x1080 frame without SSAA
    (the default scene at SSAA 1x): per-frame kernel time of each, the AO
    error of the trained network on that frame (covered pixels), and its
    error on the held-out frames (the plain pipeline's SSAO + blur on
    held-out seeded layouts and poses, evaluated here)."""
    import forkerrenderer_amd as F
    from forkerrenderer_amd import nao as NA
    if not os.path.exists(NA.DEFAULT_WEIGHTS):
        return {"error": "no trained weights at " + os.path.relpath(NA.DEFAULT_WEIGHTS, ROOT)}
    ck = NA.load_checkpoint(NA.DEFAULT_WEIGHTS)
    sc = F.Scene(gen_scene("c4"))
    I = sc.info
    ctx = F.Context(0)
    net = NA.NeuralAO(ctx, batch=1, crop=64, params=ck["state"][:NA.param_count()])
    for _ in range(warmup):
        ctx.render_scene(sc)
        net.infer()
    ctx.sync()

    def timed_frames(only):
        # frames of the plain pipeline (SSAO + blur, lighting), each followed
        # by the network and the lighting pass again, now reading the network
        # AO from its grid; HIP events only around `only` (None: every kernel)
        ctx.read_timings()
        ctx.enable_timing(True, only)
        for _ in range(steps):
            ctx.render_scene(sc)
            net.infer()
            ctx.draw_screen_space_pixels()
        ctx.sync()
        got = ctx.read_timings()
        ctx.enable_timing(False)
        return got
    # each AO stage as one span: SSAO + blur (two kernels), the network (down
    # pass, two fused kernels); the lighting pass's difference between reading
    # the AO plane and interpolating the network grid counts to the network
    plain_v = [ms for name, ms in timed_frames("ssao_blur") if name == "ssao_blur"]
    nao_v = [ms for name, ms in timed_frames("nao_conv") if name == "nao_conv"]
    light_v = [ms for name, ms in timed_frames("lighting") if name == "lighting"]
    light_plain, light_net = float(np.mean(light_v[0::2])), float(np.mean(light_v[1::2]))
    per = {}
    for name, ms in timed_frames(None):
        per[name] = per.get(name, 0.0) + ms / steps
    ctx.render_scene(sc)
    W, H = I.width, I.height
    plain = ctx.read(F.TARGET_AO).reshape(H, W)
    cover = ctx.read(F.TARGET_DEPTH).reshape(H, W) < 3e38
    net.infer()
    pred = ctx.read(F.TARGET_AO).reshape(H, W)
    err = np.abs(pred - plain)[cover]
    fixed, audited, failed = NA.audit_fallback(pred, plain, cover, 0)
    plain_ms = float(np.mean(plain_v))
    nao_span_ms = float(np.mean(nao_v))   # down pass and the two fused kernels
    nao_ms = nao_span_ms + (light_net - light_plain)
    scale = int(ck["scale"])
    net_px = -(-W // scale) * -(-H // scale)
    # per network pixel: layer 1 on the 8 real inputs, three 32-channel
    # layers, the head (multiply-adds counted as two operations)
    flop = net_px * (2 * 32 * 9 * 8 + 3 * 2 * 32 * 9 * 32 + 2 * 32)
    conv_tflops = flop / (nao_ms * 1e-3) / 1e12 if nao_ms else None
    # held-out frames, evaluated in this run against the current plain SSAO
    held = None
    if heldout_frames > 0:
        t0 = time.perf_counter()
        sset = NA.SceneSet(ctx, os.path.join(ROOT, "_scenes", "_nao"))
        held = NA.evaluate(net, ctx, sset, heldout_frames, unblurred=True)
        held["seconds"] = time.perf_counter() - t0
    net.close()
    ctx.close()
    return {"resolution": [W, H], "network_grid": [-(-W // scale), -(-H // scale)],
            "plain_ssao_blur_ms": plain_ms, "neural_ao_ms": nao_ms,
            "speedup": plain_ms / nao_ms if nao_ms else None,
            "timing": "each stage as one span of HIP events on the frame's stream (SSAO + blur: two kernels; "
                      "network: down pass and two fused convolution kernels, which leave the AO on the network "
                      "grid) plus, for the network, the lighting pass's time reading the grid minus its time "
                      "reading the AO plane; means over the frames",
            "neural_ao_span_ms": nao_span_ms, "lighting_ms": {"ao_plane": light_plain, "network_grid": light_net},
            "kernels_ms": {k: v for k, v in per.items() if k.startswith("nao") or k in ("ssao", "blur_gaussian")},
            "conv_roofline": {"bound": "mfma", "achieved": conv_tflops, "peak": 2500.0, "unit": "TFLOP/s",
                              "frac": conv_tflops / 2500.0 if conv_tflops else None,
                              "flop": flop, "note": "network FLOPs over the whole inference time (down and up "
                                                    "passes included)"},
            "frame_mae": float(err.mean()), "frame_max_err": float(err.max()),
            "frame_fallback_tiles": [failed, audited],
            "heldout_eval": held, "train_steps": int(ck["step"])}


README_DIABLO_S = 0.106988 + 3.97755   # shadow pass + forward pass, the reference's README.md:206,213


def readme_scene(stream=True):
    """The README's console-output run (README.md:190-223): obj/diablo_pose
    (5,022 faces, three 1024^2 textures) alone, forward Blinn-Phong,
    1280x800, shadow on, SSAA and SSAO off; the asset comes from the
    committed byte-for-byte fixtures.  stream: PCSS as the reference builds it
    (SOFT_SHADOW_PCSS, shadow.h:15-16), its disk samples drawn from the
    random stream for every fragment that passes the depth test -- the
    scene's default; otherwise the same counts and sizes on one fixed disk
    (`shadowfilter pcss 32 64 2.5`)."""
    from tests import assets as A
    name = "diablo_stream.scene" if stream else "diablo.scene"
    return A.scene(os.path.join(ROOT, "_scenes", "readme", name), ["diablo_pose"],
                   extra=() if stream else ("shadowfilter pcss 32 64 2.5",))


def run_readme_scene(steps, warmup, stream=True):
    """Whole frames of the README scene on the device (the scene resident),
    with the per-kernel breakdown from a separate run timed with HIP events."""
    import forkerrenderer_amd as F
    sc = F.Scene(readme_scene(stream))
    ctx = F.Context(0)
    ctx.set_tile_size(16)
    for _ in range(warmup):
        ctx.render_scene(sc)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.render_scene(sc)
    ctx.sync()
    sec = (time.perf_counter() - t0) / steps
    draws = ctx.rng_draws()
    ctx.read_timings()
    ctx.enable_timing(True, None)
    for _ in range(steps):
        ctx.render_scene(sc)
    ctx.sync()
    per = {}
    for name, ms in ctx.read_timings():
        per[name] = per.get(name, 0.0) + ms / steps
    ctx.enable_timing(False)
    ctx.close()
    ntri, _ = scene_stats(sc)
    shadow = ("PCSS as the reference builds it: 32 blocker-search and 64 PCF unit-disk samples drawn from its "
              "mt19937 stream for every fragment that passes the depth test, in raster order" if stream else
              "PCSS with the reference's sample counts and sizes (32 blocker-search / 64 PCF samples, area 2.5) "
              "on one fixed disk instead of its per-fragment random disk")
    out = {"scene": "README console-output run: obj/diablo_pose alone, forward Blinn-Phong, 1280x800, shadow on: "
                    + shadow + "; SSAA off, SSAO off", "triangles": ntri, "frames": steps,
           "ms_per_frame": sec * 1e3, "frames_per_s": 1.0 / sec, "kernel_ms_per_frame": per,
           "published_s_per_frame": README_DIABLO_S,
           "published_note": "shadow pass 0.106988 s + forward pass 3.97755 s (README.md:206,213), the "
                             "reference's default build (random-stream PCSS); hardware and build type unstated",
           "speedup_vs_published": README_DIABLO_S / sec}
    if stream:
        out["stream_draws_per_frame"] = draws
    return sc, out


def readme_cpu(sc, reps):
    """The oracle renders the README scene's whole frame (shadow pass, forward
    pass, quantise) on the current (pinned) core; best of `reps` runs."""
    import oracle as O
    I = sc.info
    models = [sc.model(i) for i in range(I.n_models)]
    mats = [sc.material(i) for i in range(I.n_materials)]
    texs = [sc.texture(i) for i in range(I.n_textures)]
    best = None
    for _ in range(reps):
        t = time.perf_counter()
        O.render_frame(models, mats, texs, I)
        dt = time.perf_counter() - t
        best = dt if best is None else min(best, dt)
    return best


def run_batched_views(reps, n_views=64):
    """The batched-views configuration: 64 views of distinct seeded scenes of
    the 1280x800 generator (2048^2 maps and shadow maps, 5x5 PCF, icosphere
    subdivision 4..8), 256^2 .. 1920x1080, deferred PBR + neural AO, as one
    batch: aggregate Mpix/s and Mtri/s of graph replays after the batch
    autotuner (tiles, rasteriser workgroups per CU, pair capacity), and the
    per-view latency of the same views rendered one at a time (render_view,
    synchronised per view)."""
    import forkerrenderer_amd as F
    from forkerrenderer_amd import nao as NA
    from forkerrenderer_amd import views as VW
    from forkerrenderer_amd.autotune import autotune_batch
    ctx = F.Context(0)
    params = NA.load_weights() if os.path.exists(NA.DEFAULT_WEIGHTS) else NA.init_params(0)
    net = NA.NeuralAO(ctx, batch=1, crop=64, params=params)
    t0 = time.perf_counter()
    specs = VW.view_specs(n_views)
    views = VW.distinct_views(ctx, specs)
    t_setup = time.perf_counter() - t0
    b = VW.Batch(ctx, views, VW.AO_NEURAL, net)
    b.render()
    b.sync()
    best, tune = autotune_batch(b)

    def timed(k):
        b.render()
        b.sync()
        t = time.perf_counter()
        for _ in range(k):
            b.render()
        b.sync()
        return (time.perf_counter() - t) / k
    sec = timed(reps)
    st = b.stats()
    # per-pass kernel time of one eager batch render
    ctx.read_timings()
    ctx.enable_timing(True)
    b.render()
    b.sync()
    per = {}
    for name, ms in ctx.read_timings():
        per[name] = per.get(name, 0.0) + ms
    ctx.enable_timing(False)
    b.close()
    lat = []
    for v in views:
        VW.render_view(ctx, v, VW.AO_NEURAL, net)
        ctx.sync()
        t = time.perf_counter()
        VW.render_view(ctx, v, VW.AO_NEURAL, net)
        ctx.sync()
        lat.append((time.perf_counter() - t) * 1e3)
    net.close()
    ctx.close()
    lat = np.array(lat)
    return {"views": n_views, "scenes": "distinct seeded scenes (view seeds 30000..), 2048^2 maps and shadow maps",
            "pixels": st["pixels"], "triangles_per_pass": st["tris_main"],
            "ms_per_batch": sec * 1e3, "mpix_per_s": st["pixels"] / sec / 1e6,
            "mtri_per_s": st["tris_main"] / sec / 1e6, "autotuned": best,
            "autotune_ms": tune, "pairs": [st["pairs_shadow"], st["pairs_main"]],
            "pair_capacity": [st["cap_shadow"], st["cap_main"]],
            "kernels_ms": per,
            "single_view_latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                                       "sum": float(lat.sum())},
            "batch_speedup_vs_sequential": float(lat.sum() / (sec * 1e3)), "setup_s": t_setup}


def host_cpu():
    """The host CPU's model name (/proc/cpuinfo) and logical core count."""
    model = platform.processor() or platform.machine()
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, os.cpu_count()


def cpu_baseline(sc, band_frac):
    """The oracle renders the whole frame -- shadow pass, geometry pass, SSAO,
    blur, lighting, quantise + SSAA -- single-threaded with the process
    pinned to one core; each pass timed.  A band estimate (the passes on a
    horizontal band of rows, scaled to all rows) is kept beside it."""
    import oracle as O
    I = sc.info
    models = [sc.model(i) for i in range(I.n_models)]
    mats = [sc.material(i) for i in range(I.n_materials)]
    texs = [sc.texture(i) for i in range(I.n_textures)]
    core = sorted(os.sched_getaffinity(0))[0]
    os.sched_setaffinity(0, {core})
    osc = O.OracleScene(models, mats, texs)
    st = O.frame_setup(I)
    W, H, SW, SH, k = st["W"], st["H"], st["SW"], st["SH"], st["k"]
    fr = st["fr"]
    lib = O.lib()
    detail = {}

    def timed(name, fn):
        t = time.perf_counter()
        r = fn()
        detail[name] = time.perf_counter() - t
        return r
    t_all = time.perf_counter()
    sd, smap = timed("shadow", lambda: O.shadow_pass(osc, st["fr_shadow"], SW, SH))
    depth, g = timed("geometry", lambda: O.geometry_pass(osc, fr, W, H))
    ao = timed("ssao", lambda: O.ssao_kernel(fr, W, H, depth, g, I.ssao))
    ao = timed("blur", lambda: O.blur_separable(ao, W, H, I.ssao.blur_weights, I.ssao.blur_radius))
    fb = timed("lighting", lambda: O.lighting_pass(fr, W, H, g, ao, smap, SW, SH))
    timed("resolve", lambda: O.ssaa_resolve(O.quantize_rgb(fb, W, H), W, H, k))
    total = time.perf_counter() - t_all
    # band estimate of the same frame (secondary): est = t_empty + (t_band - t_empty) * rows / band
    band = max(1, int(H * band_frac))
    est = detail["blur"] + detail["resolve"]

    def band_time(fn, rows, b):
        lib.orc_set_rows(0, 0)
        t = time.perf_counter(); fn(); te = time.perf_counter() - t
        y0 = rows // 2 - b // 2
        lib.orc_set_rows(y0, y0 + b)
        t = time.perf_counter(); fn(); tb = time.perf_counter() - t
        lib.orc_set_rows(0, 1 << 30)
        return te + max(tb - te, 0.0) * rows / b
    est += band_time(lambda: O.shadow_pass(osc, st["fr_shadow"], SW, SH), SH, max(1, int(SH * band_frac)))
    est += band_time(lambda: O.geometry_pass(osc, fr, W, H), H, band)
    est += band_time(lambda: O.ssao_kernel(fr, W, H, depth, g, I.ssao), H, band)
    est += band_time(lambda: O.lighting_pass(fr, W, H, g, ao, smap, SW, SH), H, band)
    return total, detail, core, est, band


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="c3")
    ap.add_argument("--tile", type=int, default=0, help="tile edge 8/16/32; 0 = autotune")
    ap.add_argument("--tiles", type=int, nargs=2, metavar=("SHADOW", "MAIN"), default=None,
                    help="per-pass tile edges instead of autotuning")
    ap.add_argument("--cpu-band", type=float, default=1.0 / 32)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-nao", action="store_true", help="skip the neural-AO vs plain SSAO comparison")
    ap.add_argument("--no-views", action="store_true", help="skip the batched uneven-views workload")
    ap.add_argument("--no-readme", action="store_true", help="skip the README's Diablo frame")
    ap.add_argument("--nao-heldout", type=int, default=200, help="held-out frames the neural-AO leg evaluates")
    ap.add_argument("--copy-calibrate", type=int, default=0, metavar="MIB",
                    help="only run a plain copy of MIB MiB (calibrates HBM byte counters) and exit")
    ap.add_argument("--readme-only", action="store_true",
                    help="only render the README's Diablo frame (both shadow variants) and print their figures")
    a = ap.parse_args()
    if a.readme_only:
        out = {}
        for stream, key in ((True, "readme_scene"), (False, "readme_scene_fixed_disk")):
            out[key] = run_readme_scene(a.steps, a.warmup, stream)[1]
        print(json.dumps(out))
        return
    if a.copy_calibrate:
        import forkerrenderer_amd as F
        ctx = F.Context(0)
        nbytes = a.copy_calibrate << 20
        ms = ctx.copy_benchmark(nbytes, 20)
        print(json.dumps({"copy_bytes": nbytes, "ms": ms, "GBps_read_plus_write": 2 * nbytes / (ms * 1e-3) / 1e9}))
        return
    if a.gpus != 1:
        print(json.dumps({"error": "runs on one GPU only: the frame path is not sharded across GPUs "
                                   "(--gpus 1)", "n_gpus": a.gpus}))
        sys.exit(2)
    import forkerrenderer_amd as F
    p = gen_scene(a.config)
    sc = F.Scene(p)
    I = sc.info
    ntri, npos = scene_stats(sc)
    print(f"[bench] scene {p}: {ntri} triangles, {npos} vertices", file=sys.stderr, flush=True)
    def dominant(per, n):
        # the kernel with an algorithmic byte model that takes most of the frame
        cands = [(float(np.sum(v)) / n, k) for k, v in per.items() if algorithmic_bytes(k, I, ntri, npos)]
        return max(cands)[1]

    sec, per, nb, sec_all, dom, dom_ms, img, tuned = run_gpu(sc, a.steps, a.warmup, a.tile, a.tiles, dominant)
    print(f"[bench] gpu {sec * 1e3:.3f} ms/frame ({sec_all * 1e3:.3f} with every kernel timed)",
          file=sys.stderr, flush=True)
    kernels = {}
    for name, v in per.items():
        if name == "ssao_blur":   # a span over the ssao and blur kernels, not a kernel
            continue
        calls = len(v) / nb
        kernels[name] = {"ms_per_frame": float(np.sum(v) / nb), "calls_per_frame": calls,
                         "ms_per_launch": float(np.mean(v))}
    ms_launch = float(np.mean(dom_ms))
    bytes_launch = algorithmic_bytes(dom, I, ntri, npos)
    achieved = bytes_launch / (ms_launch * 1e-3) / 1e9
    k = I.ssaa_k if I.ssaa_on else 1
    line = {
        "metric": METRIC, "value": 1.0 / sec, "unit": "frames/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": sec * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic (seeded scene generator)",
        "config": {"workload": f"{a.config}: {I.width}x{I.height} SSAA {k}x, deferred PBR, "
                               f"{'PCSS' if I.shadow.mode == 2 else 'PCF'} {I.shadow_map_size}^2 shadow map, "
                               f"SSAO {I.ssao.n_kernel} + blur, {ntri} triangles",
                   "internal_resolution": [I.width * k, I.height * k], "triangles": ntri, "tile": a.tile if a.tile else (tuned[0] if tuned else None), "autotune_ms": tuned[1] if tuned else None},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(dom) if a.config == "c3" else None,
                     "bytes_per_launch": bytes_launch,
                     "ms_per_launch": ms_launch},
        "kernels": kernels,
        "kernels_run": {"frames": nb, "ms_per_frame": sec_all * 1e3,
                        "note": "every kernel timed with HIP events (a separate run; the events add to the "
                                "frame time); the throughput run times only the roofline kernel"},
    }
    if not a.no_nao:
        line["neural_ao"] = run_neural_ao(max(5, a.steps // 2), a.warmup, a.nao_heldout)
    if not a.no_views:
        line["batched_views"] = run_batched_views(max(3, a.steps // 4))
    rsc = rsc_fixed = None
    if not a.no_readme:
        rsc, line["readme_scene"] = run_readme_scene(a.steps, a.warmup, True)
        rsc_fixed, line["readme_scene_fixed_disk"] = run_readme_scene(a.steps, a.warmup, False)
    cpu = None
    if not a.no_cpu:
        # after the GPU work: the process is pinned to one core from here on
        print("[bench] CPU baseline: the oracle on the whole frame, one pinned core", file=sys.stderr, flush=True)
        t_cpu, detail, core, est, band = cpu_baseline(sc, a.cpu_band)
        model, ncpu = host_cpu()
        cpu = {"value": 1.0 / t_cpu, "unit": "frames/s", "cores": 1, "kind": "port",
               "sample": "the whole frame (shadow, geometry, SSAO, blur, lighting, quantise + SSAA) rendered "
                         "once by the oracle, single-threaded, process pinned to one core",
               "seconds_per_frame": t_cpu, "per_pass_s": detail, "pinned_core": core,
               "host_cpu": model, "host_logical_cpus": ncpu,
               "band_estimate_s": est, "band_rows": band}
        for key, r in (("readme_scene", rsc), ("readme_scene_fixed_disk", rsc_fixed)):
            if r is None:
                continue
            # the README scene's whole frame by the oracle, same pinned core, best of 3
            t_r = readme_cpu(r, 3)
            line[key].update({"cpu_port_s_per_frame": t_r, "cpu_port_cores": 1,
                              "speedup_vs_cpu_port": t_r / (line[key]["ms_per_frame"] * 1e-3)})
    line["cpu_baseline"] = cpu
    line["speedup_vs_cpu"] = None if cpu is None else (1.0 / sec) / cpu["value"]
    print(json.dumps(line))


if __name__ == "__main__":
    main()
