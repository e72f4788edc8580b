"""Views and batched views through the C ABI (fgl_render_view, fgl_batch_*).

A view is one deferred frame of Render::Render without SSAA over model
instances already uploaded into a context: shadow pass, geometry pass,
ambient occlusion (none or neural), lighting, quantisation.  A Batch renders
many views of different sizes and meshes with one launch per pass (a
persistent (view, tile) queue for the rasteriser) replayed from a HIP graph;
its results equal render_view's bit for bit.

The batched-view workload (64 views, sizes log-uniform 256^2 .. 1920x1080,
aspect 1, 4:3, 16:10 or 16:9, icosphere subdivision 4..8, deferred PBR with
5x5 PCF shadows and neural AO) is built by `uneven_views`.
"""
import ctypes as C
import math

import numpy as np

from . import (FglError, Light, Material, ShadowCfg, lib, make_model_matrix, scenegen, MAX_LIGHTS,
               TARGET_AO, TARGET_DEPTH, TARGET_FRAME, TARGET_GBUFFER, TARGET_IMAGE, TARGET_SHADOW,
               TARGET_SHADOW_DEPTH, GBUF_PLANES, SHADOW_PCF_GRID)

AO_NONE, AO_NEURAL = 0, 2


class ViewDesc(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("shadow_size", C.c_int32), ("shadow_on", C.c_int32),
                ("eye", C.c_float * 3), ("look_at", C.c_float * 3), ("n_lights", C.c_int32),
                ("lights", Light * MAX_LIGHTS), ("shadow", ShadowCfg), ("n_draws", C.c_int32),
                ("models", C.POINTER(C.c_int32)), ("matrices", C.POINTER(C.c_float))]


def _bind():
    L = lib()
    if getattr(L, "_views_bound", False):
        return L
    P = C.c_void_p
    sig = {
        "fgl_render_view": (C.c_int, [P, C.POINTER(ViewDesc), C.c_int, P]),
        "fgl_batch_create": (C.c_int, [P, C.c_int, C.POINTER(ViewDesc), C.c_int, P, C.POINTER(P)]),
        "fgl_batch_destroy": (None, [P]),
        "fgl_batch_set_tiles": (C.c_int, [P, C.c_int, C.c_int]),
        "fgl_batch_set_occupancy": (C.c_int, [P, C.c_int]),
        "fgl_batch_set_capacity": (C.c_int, [P, C.c_uint32, C.c_uint32]),
        "fgl_batch_use_graph": (C.c_int, [P, C.c_int]),
        "fgl_batch_render": (C.c_int, [P]),
        "fgl_batch_sync": (C.c_int, [P, C.POINTER(C.c_int)]),
        "fgl_batch_read": (C.c_int, [P, C.c_int, C.c_int, P, C.c_size_t]),
        "fgl_batch_stats": (C.c_int, [P, C.POINTER(C.c_uint64)]),
    }
    for k, (r, a) in sig.items():
        f = getattr(L, k)
        f.restype = r
        f.argtypes = a
    L._views_bound = True
    return L


class View:
    """A view description plus the arrays it points into (kept alive)."""

    def __init__(self, width, height, eye, look_at, draws, lights, shadow=None, shadow_size=1024):
        self.models = np.ascontiguousarray([m for m, _ in draws], np.int32)
        self.mats = np.ascontiguousarray(np.stack([np.asarray(M, np.float32).reshape(16) for _, M in draws])
                                         if draws else np.zeros((0, 16), np.float32))
        d = ViewDesc()
        d.width, d.height = int(width), int(height)
        d.shadow_on = 1 if shadow is not None else 0
        d.shadow_size = int(shadow_size) if shadow is not None else 0
        d.eye[:] = [float(x) for x in eye]
        d.look_at[:] = [float(x) for x in look_at]
        d.n_lights = len(lights)
        for i, (typ, vec, col) in enumerate(lights):
            d.lights[i].type = typ
            d.lights[i].vec[:] = [float(x) for x in vec]
            d.lights[i].color[:] = [float(x) for x in col]
        if shadow is not None:
            d.shadow = shadow
        d.n_draws = len(draws)
        d.models = self.models.ctypes.data_as(C.POINTER(C.c_int32))
        d.matrices = self.mats.ctypes.data_as(C.POINTER(C.c_float))
        self.desc = d

    @property
    def pixels(self):
        return self.desc.width * self.desc.height


def render_view(ctx, view, ao_mode=AO_NONE, nao=None):
    """The single-view path: the view through the context's ForkerGL calls."""
    rc = _bind().fgl_render_view(ctx._h, C.byref(view.desc), ao_mode, nao._h if nao is not None else None)
    ctx._chk(rc, "fgl_render_view")


_TARGET_SHAPE = {TARGET_FRAME: 3, TARGET_DEPTH: 1, TARGET_GBUFFER: GBUF_PLANES, TARGET_AO: 1}


class Batch:
    def __init__(self, ctx, views, ao_mode=AO_NONE, nao=None):
        L = _bind()
        self.ctx, self.views = ctx, list(views)
        arr = (ViewDesc * len(self.views))(*[v.desc for v in self.views])
        self._arr = arr
        self._h = C.c_void_p()
        rc = L.fgl_batch_create(ctx._h, len(self.views), arr, ao_mode, nao._h if nao is not None else None,
                                C.byref(self._h))
        ctx._chk(rc, "fgl_batch_create")

    def close(self):
        if self._h:
            _bind().fgl_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tiles(self, shadow_tile, main_tile):
        self.ctx._chk(_bind().fgl_batch_set_tiles(self._h, shadow_tile, main_tile), "fgl_batch_set_tiles")

    def set_occupancy(self, wg_per_cu):
        self.ctx._chk(_bind().fgl_batch_set_occupancy(self._h, wg_per_cu), "fgl_batch_set_occupancy")

    def set_capacity(self, shadow_pairs, main_pairs):
        self.ctx._chk(_bind().fgl_batch_set_capacity(self._h, shadow_pairs, main_pairs), "fgl_batch_set_capacity")

    def use_graph(self, on=True):
        self.ctx._chk(_bind().fgl_batch_use_graph(self._h, int(on)), "fgl_batch_use_graph")

    def render(self):
        self.ctx._chk(_bind().fgl_batch_render(self._h), "fgl_batch_render")

    def sync(self):
        n = C.c_int()
        self.ctx._chk(_bind().fgl_batch_sync(self._h, C.byref(n)), "fgl_batch_sync")
        return n.value

    def stats(self):
        s = (C.c_uint64 * 10)()
        self.ctx._chk(_bind().fgl_batch_stats(self._h, s), "fgl_batch_stats")
        keys = ["pairs_shadow", "pairs_main", "cap_shadow", "cap_main", "tris_shadow", "tris_main", "pixels",
                "views", "raster_workgroups", "tiles_main"]
        return dict(zip(keys, [int(x) for x in s]))

    def read(self, i, which):
        d = self.views[i].desc
        P = d.width * d.height
        if which == TARGET_IMAGE:
            out = np.empty(3 * P, np.uint8)
        elif which in (TARGET_SHADOW, TARGET_SHADOW_DEPTH):
            out = np.empty(d.shadow_size * d.shadow_size, np.float32)
        else:
            out = np.empty(_TARGET_SHAPE[which] * P, np.float32)
        rc = _bind().fgl_batch_read(self._h, i, which, out.ctypes.data_as(C.c_void_p), out.nbytes)
        self.ctx._chk(rc, "fgl_batch_read")
        return out


# ---------------- the uneven-views workload ----------------

def tangents(v, uv, f):
    """Per-vertex tangents from UV gradients (accumulated per face, then
    Gram-Schmidt against the vertex normal direction of the sphere)."""
    p0, p1, p2 = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    t0, t1, t2 = uv[f[:, 0]], uv[f[:, 1]], uv[f[:, 2]]
    e1, e2 = p1 - p0, p2 - p0
    d1, d2 = t1 - t0, t2 - t0
    det = d1[:, 0] * d2[:, 1] - d2[:, 0] * d1[:, 1]
    det[np.abs(det) < 1e-12] = 1e-12
    T = (e1 * d2[:, 1:2] - e2 * d1[:, 1:2]) / det[:, None]
    acc = sum(scenegen.scatter_add(f[:, k], T, len(v)) for k in range(3))
    n = v / np.maximum(np.linalg.norm(v, axis=1, keepdims=True), 1e-12)
    acc -= n * np.sum(acc * n, axis=1, keepdims=True)
    ln = np.linalg.norm(acc, axis=1, keepdims=True)
    fallback = np.cross(n, np.array([0.0, 1.0, 0.0]))
    acc = np.where(ln > 1e-12, acc / np.maximum(ln, 1e-12), fallback)
    return acc.astype(np.float32)


def pbr_maps(seed, size):
    """The generator's PBR maps (host bytes): albedo uniform RGB, roughness
    round(255 U[0.05, 1]), metalness {0, 255} with probability 1/2, tangent-
    space normal map normalize(N(0, 0.2), N(0, 0.2), 1) encoded (n + 1)/2,
    B, G, R byte order for the colour maps."""
    rng = np.random.default_rng(seed)
    alb = rng.integers(0, 256, (size, size, 3), dtype=np.uint8)
    rough = np.round(255 * rng.uniform(0.05, 1.0, (size, size))).astype(np.uint8)
    metal = ((rng.random((size, size)) < 0.5) * 255).astype(np.uint8)
    nxy = rng.normal(0, 0.2, (size, size, 2))
    n = np.concatenate([nxy, np.ones((size, size, 1))], -1)
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    nrm = np.round((n + 1) / 2 * 255).astype(np.uint8)[..., ::-1].copy()   # B, G, R bytes
    return alb, rough, metal, nrm


def pbr_material_desc(ids):
    """The PBR Material record over uploaded (or oracle-indexed) maps
    [albedo, roughness, metalness, normal]."""
    m = Material()
    m.ka[:] = (1, 1, 1)
    m.kd[:] = (1, 1, 1)
    m.ks[:] = (0.5, 0.5, 0.5)
    m.albedo[:] = (1, 1, 1)
    m.roughness, m.metalness = 0.5, 0.0
    for fld in ("map_ks", "map_ke", "map_ao", "map_pbr_ke"):
        setattr(m, fld, -1)
    m.map_kd = m.map_base = ids[0]
    m.map_rough, m.map_metal, m.map_bump = ids[1], ids[2], ids[3]
    return m


def pbr_material(ctx, seed, size=512):
    """A PBR material with generated albedo / roughness / metalness / normal
    maps (repeat, bilinear)."""
    from . import FILTER_LINEAR, WRAP_REPEAT
    alb, rough, metal, nrm = pbr_maps(seed, size)
    ids = [ctx.upload_texture(t, WRAP_REPEAT, FILTER_LINEAR) for t in (alb, rough, metal, nrm)]
    m = Material()
    m.ka[:] = (1, 1, 1)
    m.kd[:] = (1, 1, 1)
    m.ks[:] = (0.5, 0.5, 0.5)
    m.albedo[:] = (1, 1, 1)
    m.roughness, m.metalness = 0.5, 0.0
    for fld in ("map_ks", "map_ke", "map_ao", "map_pbr_ke"):
        setattr(m, fld, -1)
    m.map_kd = m.map_base = ids[0]
    m.map_rough, m.map_metal, m.map_bump = ids[1], ids[2], ids[3]
    return ctx.upload_material(m)


def sphere_model(subdiv, seed, mat):
    """The generator's displaced icosphere (radius 1, 3-octave value noise of
    amplitude 0.1) as a model dict with generated tangents."""
    v, n, uv, f = scenegen.displaced_icosphere(subdiv, seed, 0.1, 1.0)
    f32 = np.float32
    return {"pos": v.astype(f32), "uv": uv.astype(f32), "nrm": n.astype(f32), "tan": tangents(v, uv, f),
            "vi": f, "ti": f, "ni": f, "face_material": np.full(len(f), mat, np.int32), "support_pbr": True}


def ground_model(mat):
    p, n, uv, f = scenegen.plane_group(4.0)
    t = np.tile(np.array([[1.0, 0.0, 0.0]], np.float32), (len(p), 1))
    return {"pos": p, "uv": uv, "nrm": n, "tan": t, "vi": f, "ti": f, "ni": f,
            "face_material": np.full(len(f), mat, np.int32), "support_pbr": True}


def upload_sphere(ctx, subdiv, seed, mat):
    return ctx.upload_model(sphere_model(subdiv, seed, mat))


def upload_ground(ctx, mat):
    return ctx.upload_model(ground_model(mat))


def pcf_grid(n=5):
    s = ShadowCfg()
    s.mode = SHADOW_PCF_GRID
    s.grid = n
    return s


def uneven_views(ctx, n_views=64, seed=5, subdivs=(4, 5, 6, 7, 8), min_px=256 * 256, max_px=1920 * 1080,
                 shadow_size=1024, tex=512):
    """The batched-view workload: per view a displaced icosphere (subdivision
    drawn from `subdivs`) on a ground plane, a random orbit camera, pixel
    count log-uniform in [min_px, max_px] with aspect 1, 4:3, 16:10 or 16:9,
    a point light with 5x5 PCF shadows.  Returns the list of View."""
    rng = np.random.default_rng(seed)
    mat = pbr_material(ctx, seed + 1000, tex)
    spheres = {s: upload_sphere(ctx, s, seed + s, mat) for s in sorted(set(subdivs))}
    ground = upload_ground(ctx, mat)
    aspects = [1.0, 4 / 3, 16 / 10, 16 / 9]
    views = []
    for i in range(n_views):
        px = math.exp(rng.uniform(math.log(min_px), math.log(max_px)))
        a = aspects[int(rng.integers(0, 4))]
        h = max(16, int(round(math.sqrt(px / a))))
        w = max(16, min(1920, int(round(h * a))))
        h = min(h, 1080)
        s = int(rng.choice(subdivs))
        r = rng.uniform(2.6, 4.0)
        el = math.radians(rng.uniform(10, 50))
        az = rng.uniform(0, 2 * math.pi)
        eye = (r * math.cos(el) * math.cos(az), r * math.sin(el), r * math.cos(el) * math.sin(az))
        draws = [(ground, make_model_matrix((0.0, -1.1, 0.0), 0.0, 3.0)),
                 (spheres[s], make_model_matrix((0.0, 0.0, 0.0), float(rng.uniform(0, 360)), 1.0))]
        lights = [(0, (2.0, 5.0, 5.0), (3.0, 3.0, 3.0))]
        views.append(View(w, h, eye, (0.0, -0.2, 0.0), draws, lights, pcf_grid(5), shadow_size))
    return views


# ---------------- the batched-views configuration: distinct seeded scenes ----------------

class ViewSpec:
    """Everything that fixes one view of the batched-views configuration: its
    seed, frame size, sphere subdivision, camera and sphere rotation.  The
    scene itself (sphere, ground, maps) follows from the seed."""

    def __init__(self, seed, width, height, subdiv, eye, look_at, roty, tex, shadow_size):
        self.seed, self.width, self.height, self.subdiv = seed, width, height, subdiv
        self.eye, self.look_at, self.roty, self.tex, self.shadow_size = eye, look_at, roty, tex, shadow_size
        self.light = (0, (2.0, 5.0, 5.0), (3.0, 3.0, 3.0))

    def matrices(self):
        return (make_model_matrix((0.0, -1.1, 0.0), 0.0, 3.0), make_model_matrix((0.0, 0.0, 0.0), self.roty, 1.0))


def view_specs(n_views=64, seed=30000, subdivs=(4, 5, 6, 7, 8), min_px=256 * 256, max_px=1920 * 1080, tex=2048,
               shadow_size=2048):
    """The batched-views configuration: view i has seed seed + i; pixel count
    log-uniform in [min_px, max_px] with aspect 1, 4:3, 16:10 or 16:9;
    icosphere subdivision drawn from `subdivs`; a random orbit camera."""
    out = []
    aspects = [1.0, 4 / 3, 16 / 10, 16 / 9]
    for i in range(n_views):
        rng = np.random.default_rng(seed + i)
        px = math.exp(rng.uniform(math.log(min_px), math.log(max_px)))
        a = aspects[int(rng.integers(0, 4))]
        h = max(16, int(round(math.sqrt(px / a))))
        w = max(16, min(1920, int(round(h * a))))
        h = min(h, 1080)
        sd = int(rng.choice(subdivs))
        r = rng.uniform(2.6, 4.0)
        el = math.radians(rng.uniform(10, 50))
        az = rng.uniform(0, 2 * math.pi)
        eye = (r * math.cos(el) * math.cos(az), r * math.sin(el), r * math.cos(el) * math.sin(az))
        out.append(ViewSpec(seed + i, w, h, sd, eye, (0.0, -0.2, 0.0), float(rng.uniform(0, 360)), tex, shadow_size))
    return out


def scene_host(spec):
    """Host arrays of a view's scene, material index 0 and texture indices
    0..3 (albedo, roughness, metalness, normal; repeat, bilinear)."""
    return {"sphere": sphere_model(spec.subdiv, spec.seed, 0), "ground": ground_model(0),
            "maps": pbr_maps(spec.seed + 1000, spec.tex)}


def distinct_views(ctx, specs, workers=8):
    """Uploads every view's own scene (the host arrays are generated in a
    thread pool) and returns the View list."""
    from concurrent.futures import ThreadPoolExecutor
    from . import FILTER_LINEAR, WRAP_REPEAT
    views = []
    with ThreadPoolExecutor(max_workers=workers) as ex:
        for spec, host in zip(specs, ex.map(scene_host, specs)):
            ids = [ctx.upload_texture(t, WRAP_REPEAT, FILTER_LINEAR) for t in host["maps"]]
            mat = ctx.upload_material(pbr_material_desc(ids))
            for k in ("sphere", "ground"):
                host[k]["face_material"] = np.full(len(host[k]["vi"]), mat, np.int32)
            sphere = ctx.upload_model(host["sphere"])
            ground = ctx.upload_model(host["ground"])
            Mg, Ms = spec.matrices()
            views.append(View(spec.width, spec.height, spec.eye, spec.look_at, [(ground, Mg), (sphere, Ms)],
                              [spec.light], pcf_grid(5), spec.shadow_size))
    return views
