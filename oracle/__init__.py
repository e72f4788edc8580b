"""CPU oracle for the ForkerRenderer raster/shading path.

TEST INFRASTRUCTURE ONLY: the checker that the HIP path is compared with.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it;
the product package never imports it.

liboracle.so (oracle/fgl_oracle.c) restates the reference's arithmetic; this
module binds it with ctypes and drives it pass by pass the way Render::Render
does (src/render.cpp:40-58).

Parity pinning of each oracle function (how it is checked independently of
retyping it) is listed in the test files tests/test_oracle_*.py.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_DIR, "liboracle.so")


def build():
    subprocess.run(["make", "-s", "-C", _DIR], check=True)
    return LIB_PATH


class Texture(C.Structure):
    _fields_ = [("w", C.c_int), ("h", C.c_int), ("bpp", C.c_int), ("data", C.c_void_p),
                ("wrap", C.c_int), ("filter", C.c_int)]


class Material(C.Structure):
    _fields_ = [("ka", C.c_float * 3), ("kd", C.c_float * 3), ("ks", C.c_float * 3), ("ke", C.c_float * 3),
                ("map_kd", C.c_int), ("map_ks", C.c_int), ("map_ke", C.c_int), ("map_bump", C.c_int),
                ("pbr_ke", C.c_float * 3), ("albedo", C.c_float * 3), ("roughness", C.c_float),
                ("metalness", C.c_float), ("map_base", C.c_int), ("map_rough", C.c_int),
                ("map_metal", C.c_int), ("map_ao", C.c_int), ("map_pbr_ke", C.c_int)]


class Draw(C.Structure):
    _fields_ = [("n_faces", C.c_int), ("vi", C.c_void_p), ("ti", C.c_void_p), ("ni", C.c_void_p),
                ("pos", C.c_void_p), ("uv", C.c_void_p), ("nrm", C.c_void_p), ("tan", C.c_void_p),
                ("has_tangents", C.c_int), ("support_pbr", C.c_int), ("material", C.c_int),
                ("model", C.c_float * 16), ("normal", C.c_float * 9)]


class Light(C.Structure):
    _fields_ = [("type", C.c_int), ("vec", C.c_float * 3), ("color", C.c_float * 3)]


class ShadowCfg(C.Structure):
    _fields_ = [("mode", C.c_int), ("grid", C.c_int), ("n_blocker", C.c_int), ("n_pcf", C.c_int),
                ("blocker_disk", C.c_void_p), ("pcf_disk", C.c_void_p), ("pcf_filter", C.c_float),
                ("blocker_filter", C.c_double), ("pcss_filter", C.c_double), ("area_light", C.c_float)]


class Frame(C.Structure):
    _fields_ = [("view", C.c_float * 16), ("proj", C.c_float * 16), ("light_space", C.c_float * 16),
                ("viewport", C.c_float * 16), ("shadow_viewport", C.c_float * 16), ("eye", C.c_float * 3),
                ("shadow_on", C.c_int), ("shadow", ShadowCfg), ("n_lights", C.c_int),
                ("lights", Light * 4)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        sig = {
            "orc_mat_mul": (None, [P, P, P]),
            "orc_make_model_matrix": (None, [C.c_float] * 5 + [P]),
            "orc_make_lookat": (None, [P, P, P, P]),
            "orc_make_perspective": (None, [C.c_float] * 4 + [P]),
            "orc_make_ortho": (None, [C.c_float] * 6 + [P]),
            "orc_make_normal_matrix": (None, [P, P]),
            "orc_make_viewport": (None, [C.c_int] * 4 + [P]),
            "orc_trunc_int": (C.c_int, [C.c_float]),
            "orc_barycentric": (C.c_int, [C.c_int] * 8 + [P]),
            "orc_texture_sample": (None, [C.POINTER(Texture), C.c_float, C.c_float, P]),
            "orc_texture_sample_float": (C.c_float, [C.POINTER(Texture), C.c_float, C.c_float]),
            "orc_shadow_pass": (None, [P, C.c_int, C.POINTER(Frame), C.c_int, C.c_int, P, P]),
            "orc_geometry_pass": (None, [P, C.c_int, P, P, C.POINTER(Frame), C.c_int, C.c_int, P, P]),
            "orc_forward_pass": (None, [P, C.c_int, P, P, C.POINTER(Frame), C.c_int, C.c_int, P, C.c_int,
                                        C.c_int, P, P]),
            "orc_lighting_pass": (None, [C.POINTER(Frame), C.c_int, C.c_int, P, P, P, C.c_int, C.c_int, P]),
            "orc_shadow_visibility": (C.c_float, [C.POINTER(Frame), P, C.c_int, C.c_int, P, P, P]),
            "orc_shade_blinn_phong": (None, [P, P, P, C.c_float, P, P, P, P, C.c_int, P]),
            "orc_shade_pbr": (None, [P, P, P, P, C.c_float, P, P, P, P, C.c_int, P]),
            "orc_ssao_kernel": (None, [C.POINTER(Frame), C.c_int, C.c_int, P, P, P, P, C.c_int, P,
                                       C.c_float, C.c_float, C.c_float, P]),
            "orc_ssao_reference": (None, [C.POINTER(Frame), C.c_int, C.c_int, P, P, P, P]),
            "orc_rng_seed": (None, [C.c_uint]),
            "orc_rng_draws": (C.c_ulonglong, []),
            "orc_rng_raw": (None, [C.c_int, P]),
            "orc_rng_uniform": (None, [C.c_int, P]),
            "orc_rng_sphere": (None, [C.c_int, P]),
            "orc_rng_disk": (None, [C.c_int, P]),
            "orc_rng_hemisphere": (None, [P, P]),
            "orc_blur_recursive": (None, [P, C.c_int, C.c_int]),
            "orc_blur_separable": (None, [P, C.c_int, C.c_int, P, C.c_int, P]),
            "orc_quantize_rgb": (None, [P, C.c_int, C.c_int, P]),
            "orc_quantize_grey": (None, [P, C.c_int, C.c_int, P]),
            "orc_ssaa_resolve": (None, [P, C.c_int, C.c_int, C.c_int, P]),
        }
        for k, (r, a) in sig.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


# ---- matrices ----

def mat_mul(a, b):
    out = np.zeros(16, np.float32)
    lib().orc_mat_mul(_p(_f32(a).reshape(16)), _p(_f32(b).reshape(16)), _p(out))
    return out.reshape(4, 4)


def make_model_matrix(t, roty, scale):
    out = np.zeros(16, np.float32)
    lib().orc_make_model_matrix(float(t[0]), float(t[1]), float(t[2]), float(roty), float(scale), _p(out))
    return out.reshape(4, 4)


def make_lookat(eye, center, up=(0.0, 1.0, 0.0)):
    out = np.zeros(16, np.float32)
    lib().orc_make_lookat(_p(_f32(eye)), _p(_f32(center)), _p(_f32(up)), _p(out))
    return out.reshape(4, 4)


def make_perspective(fov, aspect, n, f):
    out = np.zeros(16, np.float32)
    lib().orc_make_perspective(float(fov), float(aspect), float(n), float(f), _p(out))
    return out.reshape(4, 4)


def make_ortho(l, r, b, t, n, f):
    out = np.zeros(16, np.float32)
    lib().orc_make_ortho(float(l), float(r), float(b), float(t), float(n), float(f), _p(out))
    return out.reshape(4, 4)


def make_normal_matrix(m):
    out = np.zeros(9, np.float32)
    lib().orc_make_normal_matrix(_p(_f32(m).reshape(16)), _p(out))
    return out.reshape(3, 3)


def make_viewport(x, y, w, h):
    out = np.zeros(16, np.float32)
    lib().orc_make_viewport(x, y, w, h, _p(out))
    return out.reshape(4, 4)


def barycentric(a, b, c, p):
    out = np.zeros(3, np.float32)
    ok = lib().orc_barycentric(int(a[0]), int(a[1]), int(b[0]), int(b[1]), int(c[0]), int(c[1]),
                               int(p[0]), int(p[1]), _p(out))
    return (out if ok else None)


# ---- scene data -> oracle structures ----

class OracleScene:
    """Holds numpy arrays alive and the ctypes structures pointing at them.

    models: list of (model_dict, model_matrix 4x4); model_dict as produced by
    forkerrenderer_amd.Scene.model() or oracle.loader.  materials: list of
    Material-like objects with the fgl_material fields.  textures: list of
    dicts {bytes (h,w[,bpp]) uint8 B,G,R, wrap, filter}.
    """

    def __init__(self, models, materials, textures):
        self._keep = []
        self.tex = (Texture * max(1, len(textures)))()
        for i, t in enumerate(textures):
            b = np.ascontiguousarray(t["bytes"], dtype=np.uint8)
            self._keep.append(b)
            bpp = 1 if b.ndim == 2 else b.shape[2]
            self.tex[i] = Texture(b.shape[1], b.shape[0], bpp, b.ctypes.data, int(t["wrap"]), int(t["filter"]))
        self.mats = (Material * max(1, len(materials)))()
        for i, m in enumerate(materials):
            C.memmove(C.byref(self.mats[i]), C.byref(m), C.sizeof(Material))
        draws = []
        for md, mm in models:
            pos, uv, nrm = _f32(md["pos"]), _f32(md["uv"]), _f32(md["nrm"])
            tan = None if md.get("tan") is None else _f32(md["tan"])
            vi, ti, ni = _i32(md["vi"]).reshape(-1, 3), _i32(md["ti"]).reshape(-1, 3), _i32(md["ni"]).reshape(-1, 3)
            fm = _i32(md["face_material"])
            normal = make_normal_matrix(mm)
            # one Mesh::Draw per run of faces sharing a material (mesh order kept)
            start = 0
            nf = len(fm)
            while start < nf:
                end = start + 1
                while end < nf and fm[end] == fm[start]:
                    end += 1
                sv, st, sn = (np.ascontiguousarray(x[start:end]) for x in (vi, ti, ni))
                self._keep += [sv, st, sn, pos, uv, nrm] + ([tan] if tan is not None else [])
                d = Draw()
                d.n_faces = end - start
                d.vi, d.ti, d.ni = sv.ctypes.data, st.ctypes.data, sn.ctypes.data
                d.pos, d.uv, d.nrm = pos.ctypes.data, uv.ctypes.data, nrm.ctypes.data
                d.tan = tan.ctypes.data if tan is not None else None
                d.has_tangents = int(tan is not None)
                d.support_pbr = int(bool(md["support_pbr"]))
                d.material = int(fm[start])
                d.model[:] = _f32(mm).reshape(16).tolist()
                d.normal[:] = normal.reshape(9).tolist()
                draws.append(d)
                start = end
        self.n_draws = len(draws)
        self.draws = (Draw * max(1, len(draws)))(*draws)


def make_frame(view, proj, light_space, viewport, shadow_viewport, eye, shadow_on, shadow_cfg, lights):
    """shadow_cfg: forkerrenderer_amd.ShadowCfg-like (fields mode, grid, n_blocker,
    n_pcf, blocker_disk, pcf_disk, pcf_filter, blocker_filter, pcss_filter, area_light)."""
    fr = Frame()
    fr.view[:] = _f32(view).reshape(16).tolist()
    fr.proj[:] = _f32(proj).reshape(16).tolist()
    fr.light_space[:] = _f32(light_space).reshape(16).tolist()
    fr.viewport[:] = _f32(viewport).reshape(16).tolist()
    fr.shadow_viewport[:] = _f32(shadow_viewport).reshape(16).tolist()
    fr.eye[:] = [float(x) for x in eye]
    fr.shadow_on = int(shadow_on)
    keep = []
    if shadow_cfg is not None:
        bd = np.array(list(shadow_cfg.blocker_disk), np.float32)
        pd = np.array(list(shadow_cfg.pcf_disk), np.float32)
        keep += [bd, pd]
        fr.shadow.mode, fr.shadow.grid = shadow_cfg.mode, shadow_cfg.grid
        fr.shadow.n_blocker, fr.shadow.n_pcf = shadow_cfg.n_blocker, shadow_cfg.n_pcf
        fr.shadow.blocker_disk, fr.shadow.pcf_disk = bd.ctypes.data, pd.ctypes.data
        fr.shadow.pcf_filter = shadow_cfg.pcf_filter
        fr.shadow.blocker_filter = shadow_cfg.blocker_filter
        fr.shadow.pcss_filter = shadow_cfg.pcss_filter
        fr.shadow.area_light = shadow_cfg.area_light
    fr.n_lights = len(lights)
    for i, (typ, vec, col) in enumerate(lights):
        fr.lights[i].type = typ
        fr.lights[i].vec[:] = [float(x) for x in vec]
        fr.lights[i].color[:] = [float(x) for x in col]
    fr._keep = keep
    return fr


def shadow_pass(sc, fr, W, H):
    depth = np.empty(W * H, np.float32)
    shadow = np.empty(W * H, np.float32)
    lib().orc_shadow_pass(C.cast(sc.draws, C.c_void_p), sc.n_draws, C.byref(fr), W, H, _p(depth), _p(shadow))
    return depth, shadow


def _planes(n, P):
    arrs = [np.empty(P, np.float32) for _ in range(n)]
    ptrs = (C.c_void_p * n)(*[a.ctypes.data for a in arrs])
    return arrs, ptrs


def geometry_pass(sc, fr, W, H):
    depth = np.empty(W * H, np.float32)
    g, gp = _planes(19, W * H)
    lib().orc_geometry_pass(C.cast(sc.draws, C.c_void_p), sc.n_draws, C.cast(sc.mats, C.c_void_p),
                            C.cast(sc.tex, C.c_void_p), C.byref(fr), W, H, _p(depth), C.cast(gp, C.c_void_p))
    return depth, np.stack(g)


def forward_pass(sc, fr, W, H, shadow_map, SW, SH):
    depth = np.empty(W * H, np.float32)
    fb, fp = _planes(3, W * H)
    sm = _f32(shadow_map) if shadow_map is not None else np.zeros(1, np.float32)
    lib().orc_forward_pass(C.cast(sc.draws, C.c_void_p), sc.n_draws, C.cast(sc.mats, C.c_void_p),
                           C.cast(sc.tex, C.c_void_p), C.byref(fr), W, H, _p(sm), SW, SH, _p(depth),
                           C.cast(fp, C.c_void_p))
    return depth, np.stack(fb)


def lighting_pass(fr, W, H, gbuf, ao, shadow_map, SW, SH):
    g = [np.ascontiguousarray(gbuf[i], np.float32) for i in range(19)]
    gp = (C.c_void_p * 19)(*[a.ctypes.data for a in g])
    fb, fp = _planes(3, W * H)
    aoa = _f32(ao)
    sm = _f32(shadow_map) if shadow_map is not None else np.zeros(1, np.float32)
    lib().orc_lighting_pass(C.byref(fr), W, H, C.cast(gp, C.c_void_p), _p(aoa), _p(sm), SW, SH,
                            C.cast(fp, C.c_void_p))
    return np.stack(fb)


def ssao_kernel(fr, W, H, depth, gbuf, cfg):
    n = [np.ascontiguousarray(gbuf[i], np.float32) for i in range(3)]
    p = [np.ascontiguousarray(gbuf[3 + i], np.float32) for i in range(3)]
    npp = (C.c_void_p * 3)(*[a.ctypes.data for a in n])
    ppp = (C.c_void_p * 3)(*[a.ctypes.data for a in p])
    k = np.array(list(cfg.kernel), np.float32)
    nz = np.array(list(cfg.noise), np.float32)
    ao = np.empty(W * H, np.float32)
    d = _f32(depth)
    lib().orc_ssao_kernel(C.byref(fr), W, H, _p(d), C.cast(npp, C.c_void_p), C.cast(ppp, C.c_void_p), _p(k),
                          cfg.n_kernel, _p(nz), cfg.radius, cfg.bias, cfg.range, _p(ao))
    return ao


def ssao_reference(fr, W, H, depth, gbuf):
    """Render::DoSSAO with the reference's random stream (continues the
    oracle's stream; see rng_seed)."""
    n = [np.ascontiguousarray(gbuf[i], np.float32) for i in range(3)]
    p = [np.ascontiguousarray(gbuf[3 + i], np.float32) for i in range(3)]
    npp = (C.c_void_p * 3)(*[a.ctypes.data for a in n])
    ppp = (C.c_void_p * 3)(*[a.ctypes.data for a in p])
    ao = np.empty(W * H, np.float32)
    d = _f32(depth)
    lib().orc_ssao_reference(C.byref(fr), W, H, _p(d), C.cast(npp, C.c_void_p), C.cast(ppp, C.c_void_p), _p(ao))
    return ao


# ---- the reference's random stream (oracle/ref_rng.cpp) ----

def rng_seed(seed=5489):
    lib().orc_rng_seed(seed)


def rng_draws():
    return int(lib().orc_rng_draws())


def rng_raw(n):
    out = np.empty(n, np.uint32)
    lib().orc_rng_raw(n, _p(out))
    return out


def rng_uniform(n):
    out = np.empty(n, np.float32)
    lib().orc_rng_uniform(n, _p(out))
    return out


def rng_sphere(n):
    out = np.empty((n, 3), np.float32)
    lib().orc_rng_sphere(n, _p(out))
    return out


def rng_disk(n):
    out = np.empty((n, 2), np.float32)
    lib().orc_rng_disk(n, _p(out))
    return out


def blur_recursive(ao, W, H):
    a = _f32(ao).copy()
    lib().orc_blur_recursive(_p(a), W, H)
    return a


def blur_separable(ao, W, H, weights, R):
    a = _f32(ao).copy()
    tmp = np.empty_like(a)
    w = _f32(list(weights)[:9])
    lib().orc_blur_separable(_p(a), W, H, _p(w), R, _p(tmp))
    return a


def quantize_rgb(fb, W, H):
    f = [np.ascontiguousarray(fb[i], np.float32) for i in range(3)]
    fp = (C.c_void_p * 3)(*[a.ctypes.data for a in f])
    out = np.empty(W * H * 3, np.uint8)
    lib().orc_quantize_rgb(C.cast(fp, C.c_void_p), W, H, _p(out))
    return out


def quantize_grey(plane, W, H):
    """Buffer1f::GenerateImage (src/buffer.cpp:20-32) on one float plane."""
    a = np.ascontiguousarray(plane, np.float32)
    out = np.empty(W * H, np.uint8)
    lib().orc_quantize_grey(_p(a), W, H, _p(out))
    return out


def ssaa_resolve(img, W, H, k):
    out = np.empty((W // k) * (H // k) * 3, np.uint8)
    lib().orc_ssaa_resolve(_p(np.ascontiguousarray(img, np.uint8)), W, H, k, _p(out))
    return out


def frame_setup(info, shadow_cfg=None):
    """The per-frame state Render::Render sets up (src/render.cpp:60-131):
    sizes, viewports, the light-space matrix of lights[0] (ortho +-3*ratio,
    0.1..20, looking at the origin) and the camera's view / projection, as
    oracle Frame records for the shadow pass and for the main passes."""
    k = info.ssaa_k if info.ssaa_on else 1
    W, H = info.width * k, info.height * k
    ratio = np.float32(info.width) / np.float32(info.height)
    S = info.shadow_map_size
    SW, SH = (S, S) if S > 0 else (W, H)
    lights = [(info.lights[i].type, list(info.lights[i].vec), list(info.lights[i].color))
              for i in range(info.n_lights)]
    vp = make_viewport(0, 0, W, H)
    svp = make_viewport(0, 0, SW, SH)
    shadow_cfg = shadow_cfg if shadow_cfg is not None else info.shadow
    L = np.eye(4, dtype=np.float32)
    fr_shadow = None
    if info.shadow_on:
        V = make_lookat(lights[0][1], (0, 0, 0))
        P = make_ortho(np.float32(-3.0) * ratio, np.float32(3.0) * ratio, -3.0, 3.0, 0.1, 20.0)
        L = mat_mul(P, V)
        fr_shadow = make_frame(np.eye(4), np.eye(4), L, vp, svp, info.eye, 1, shadow_cfg, lights)
    V = make_lookat(list(info.eye), list(info.look_at))
    if info.camera_ortho:
        P = make_ortho(np.float32(-1.0) * ratio, np.float32(1.0) * ratio, -1.0, 1.0, 0.01, 20.0)
    else:
        P = make_perspective(45.0, ratio, 0.01, 20.0)
    fr = make_frame(V, P, L, vp, svp, info.eye, info.shadow_on, shadow_cfg, lights)
    return dict(k=k, W=W, H=H, SW=SW, SH=SH, lights=lights, view=V, proj=P, light_space=L, viewport=vp,
                shadow_viewport=svp, shadow_cfg=shadow_cfg, fr_shadow=fr_shadow, fr=fr)


def render_frame(models, materials, textures, info, shadow_cfg=None, ssao_cfg=None, multilight_lights=None,
                 rng_start=5489):
    """Render::Render on the oracle (src/render.cpp:40-58), with the same
    extension settings the product reads from the .scene file.  `info` is a
    forkerrenderer_amd.SceneInfo-like object.  Returns a dict of buffers."""
    sc = OracleScene(models, materials, textures)
    rng_seed(rng_start)   # one Render::Render run: the stream starts fresh
    st = frame_setup(info, shadow_cfg)
    k, W, H, SW, SH = st["k"], st["W"], st["H"], st["SW"], st["SH"]
    out = {}
    ssao_cfg = ssao_cfg if ssao_cfg is not None else info.ssao
    smap = None
    if info.shadow_on:
        sd, smap = shadow_pass(sc, st["fr_shadow"], SW, SH)
        out["shadow_depth"], out["shadow"] = sd, smap
    fr = st["fr"]
    if not info.deferred:
        depth, fb = forward_pass(sc, fr, W, H, smap, SW, SH)
        out["depth"] = depth
    else:
        depth, g = geometry_pass(sc, fr, W, H)
        out["depth"], out["gbuf"] = depth, g
        ao = np.ones(W * H, np.float32)
        if info.ssao_on:
            if getattr(ssao_cfg, "stochastic", 0):
                ao = ssao_reference(fr, W, H, depth, g)
            else:
                ao = ssao_kernel(fr, W, H, depth, g, ssao_cfg)
            if ssao_cfg.blur == 1:
                ao = blur_recursive(ao, W, H)
            elif ssao_cfg.blur == 2:
                ao = blur_separable(ao, W, H, ssao_cfg.blur_weights, ssao_cfg.blur_radius)
        out["ao"] = ao
        fb = lighting_pass(fr, W, H, g, ao, smap, SW, SH)
    out["fb"] = fb
    q = quantize_rgb(fb, W, H)
    out["image"] = ssaa_resolve(q, W, H, k) if k > 1 else q
    return out
