"""forkerrenderer_amd -- MI355X (gfx950) raster/shading path with the
reference ForkerRenderer's ForkerGL interface.

Python mirror of the C ABI in include/fgl.h (ctypes over the in-tree
libforkergl.so).  The library must have been built (forkerrenderer_amd/build.py
or __graft_entry__.build()); there is no CPU fallback -- importing the
bindings without the library raises.
"""
import ctypes as C
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "_lib", "libforkergl.so")

FGL_OK = 0
PASS_FORWARD, PASS_GEOMETRY, PASS_LIGHTING, PASS_SHADOW = 0, 1, 2, 3
WRAP_NONE, WRAP_REPEAT, WRAP_MIRROR, WRAP_CLAMP = 0, 1, 2, 3
FILTER_NEAREST, FILTER_LINEAR = 0, 1
MAT_VIEW, MAT_PROJ, MAT_LIGHT_SPACE = 0, 1, 2
SHADOW_HARD, SHADOW_PCF_DISK, SHADOW_PCSS, SHADOW_PCF_GRID = 0, 1, 2, 3
SHADOW_REF_PCF, SHADOW_REF_PCSS = 4, 5
TARGET_FRAME, TARGET_DEPTH, TARGET_SHADOW, TARGET_GBUFFER, TARGET_AO, TARGET_IMAGE, TARGET_SHADOW_DEPTH = range(7)
MAX_LIGHTS, MAX_DISK = 4, 64
GBUF_PLANES = 19


# status codes (include/fgl.h)
OK, ERR_INVALID, ERR_DEVICE, ERR_IO, ERR_NOMEM, ERR_CAPACITY = 0, -1, -2, -3, -4, -5


class FglError(RuntimeError):
    pass


class Material(C.Structure):
    _fields_ = [("ka", C.c_float * 3), ("kd", C.c_float * 3), ("ks", C.c_float * 3), ("ke", C.c_float * 3),
                ("map_kd", C.c_int32), ("map_ks", C.c_int32), ("map_ke", C.c_int32), ("map_bump", C.c_int32),
                ("pbr_ke", C.c_float * 3), ("albedo", C.c_float * 3), ("roughness", C.c_float),
                ("metalness", C.c_float), ("map_base", C.c_int32), ("map_rough", C.c_int32),
                ("map_metal", C.c_int32), ("map_ao", C.c_int32), ("map_pbr_ke", C.c_int32)]


class ModelDesc(C.Structure):
    _fields_ = [("n_pos", C.c_int32), ("n_uv", C.c_int32), ("n_nrm", C.c_int32), ("n_faces", C.c_int32),
                ("pos", C.POINTER(C.c_float)), ("uv", C.POINTER(C.c_float)), ("nrm", C.POINTER(C.c_float)),
                ("tan", C.POINTER(C.c_float)), ("vi", C.POINTER(C.c_int32)), ("ti", C.POINTER(C.c_int32)),
                ("ni", C.POINTER(C.c_int32)), ("face_material", C.POINTER(C.c_int32)),
                ("support_pbr", C.c_int32)]


class Light(C.Structure):
    _fields_ = [("type", C.c_int32), ("vec", C.c_float * 3), ("color", C.c_float * 3)]


class ShadowCfg(C.Structure):
    _fields_ = [("mode", C.c_int32), ("grid", C.c_int32), ("n_blocker", C.c_int32), ("n_pcf", C.c_int32),
                ("blocker_disk", C.c_float * (2 * MAX_DISK)), ("pcf_disk", C.c_float * (2 * MAX_DISK)),
                ("pcf_filter", C.c_float), ("blocker_filter", C.c_double), ("pcss_filter", C.c_double),
                ("area_light", C.c_float)]


class SsaoCfg(C.Structure):
    _fields_ = [("n_kernel", C.c_int32), ("kernel", C.c_float * 192), ("noise", C.c_float * 48),
                ("radius", C.c_float), ("bias", C.c_float), ("range", C.c_float), ("blur", C.c_int32),
                ("blur_radius", C.c_int32), ("blur_weights", C.c_float * 9), ("stochastic", C.c_int32)]


class SceneInfo(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("ssaa_on", C.c_int32), ("ssaa_k", C.c_int32),
                ("ssao_on", C.c_int32), ("shadow_on", C.c_int32), ("deferred", C.c_int32),
                ("camera_ortho", C.c_int32), ("eye", C.c_float * 3), ("look_at", C.c_float * 3),
                ("n_lights", C.c_int32), ("lights", Light * MAX_LIGHTS), ("n_models", C.c_int32),
                ("n_materials", C.c_int32), ("n_textures", C.c_int32), ("shadow_map_size", C.c_int32),
                ("shadow", ShadowCfg), ("ssao", SsaoCfg)]


_lib_handle = None


def lib():
    """The loaded libforkergl.so; raises if it has not been built."""
    global _lib_handle
    if _lib_handle is None:
        if not os.path.exists(LIB_PATH):
            raise FglError(f"{LIB_PATH} is missing: build it with forkerrenderer_amd/build.py")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        sigs = {
            "fgl_ctx_create": (C.c_int, [C.c_int, C.POINTER(P)]),
            "fgl_ctx_destroy": (None, [P]),
            "fgl_last_error": (C.c_char_p, [P]),
            "fgl_sync": (C.c_int, [P]),
            "fgl_set_tile_size": (C.c_int, [P, C.c_int]),
            "fgl_set_pass_tile_size": (C.c_int, [P, C.c_int, C.c_int]),
            "fgl_get_pass_tile_size": (C.c_int, [P, C.c_int, C.POINTER(C.c_int)]),
            "fgl_upload_texture": (C.c_int, [P, C.c_int, C.c_int, C.c_int, P, C.c_int, C.c_int, C.POINTER(C.c_int)]),
            "fgl_upload_material": (C.c_int, [P, C.POINTER(Material), C.POINTER(C.c_int)]),
            "fgl_upload_model": (C.c_int, [P, C.POINTER(ModelDesc), C.POINTER(C.c_int)]),
            "fgl_init_targets": (C.c_int, [P, C.c_int, C.c_int]),
            "fgl_init_buffer": (C.c_int, [P, C.c_int, C.c_int, C.c_int]),
            "fgl_init_shadow_buffer": (C.c_int, [P, C.c_int, C.c_int]),
            "fgl_clear_color": (C.c_int, [P, C.c_float, C.c_float, C.c_float]),
            "fgl_set_viewport": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
            "fgl_set_matrix": (C.c_int, [P, C.c_int, P]),
            "fgl_set_eye": (C.c_int, [P, P]),
            "fgl_set_lights": (C.c_int, [P, C.c_int, C.POINTER(Light)]),
            "fgl_set_shadow": (C.c_int, [P, C.c_int, C.POINTER(ShadowCfg)]),
            "fgl_begin_pass": (C.c_int, [P, C.c_int]),
            "fgl_draw_model": (C.c_int, [P, C.c_int, P]),
            "fgl_end_pass": (C.c_int, [P]),
            "fgl_draw_screen_space_pixels": (C.c_int, [P]),
            "fgl_ssao": (C.c_int, [P, C.POINTER(SsaoCfg)]),
            "fgl_resolve": (C.c_int, [P, C.c_int]),
            "fgl_rng_seed": (C.c_int, [P, C.c_uint32]),
            "fgl_rng_draws": (C.c_int, [P, C.POINTER(C.c_uint64)]),
            "fgl_rng_vectors": (C.c_int, [P, C.c_int, C.c_uint32, P]),
            "fgl_read_target": (C.c_int, [P, C.c_int, P, C.c_size_t]),
            "fgl_write_target": (C.c_int, [P, C.c_int, P, C.c_size_t]),
            "fgl_target_bytes": (C.c_int, [P, C.c_int, C.POINTER(C.c_size_t)]),
            "fgl_enable_timing": (C.c_int, [P, C.c_int]),
            "fgl_time_only": (C.c_int, [P, C.c_char_p]),
            "fgl_sort_pairs": (C.c_int, [P, P, P, C.c_size_t, C.c_int]),
            "fgl_check_recip": (C.c_int, [P, P, P, P]),
            "fgl_read_timings": (C.c_int, [P, P, P, C.c_int, C.POINTER(C.c_int)]),
            "fgl_scene_load": (C.c_int, [C.c_char_p, C.POINTER(P), C.c_char_p, C.c_size_t]),
            "fgl_scene_load_with": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.POINTER(P), C.c_char_p, C.c_size_t]),
            "fgl_scene_upload": (C.c_int, [P, P, P]),
            "fgl_draw_model_normal": (C.c_int, [P, C.c_int, P, P]),
            "fgl_make_normal_matrix": (None, [P, P]),
            "fgl_scene_free": (None, [P]),
            "fgl_scene_info_get": (C.c_int, [P, C.POINTER(SceneInfo)]),
            "fgl_scene_model": (C.c_int, [P, C.c_int, C.POINTER(ModelDesc), P]),
            "fgl_scene_material": (C.c_int, [P, C.c_int, C.POINTER(Material)]),
            "fgl_scene_texture": (C.c_int, [P, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                            C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(P)]),
            "fgl_render_scene": (C.c_int, [P, P]),
            "fgl_make_model_matrix": (None, [C.c_float] * 5 + [P]),
            "fgl_make_lookat": (None, [P, P, P, P]),
            "fgl_make_perspective": (None, [C.c_float] * 4 + [P]),
            "fgl_make_ortho": (None, [C.c_float] * 6 + [P]),
            "fgl_mat_mul": (None, [P, P, P]),
            "fgl_tga_write": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, P, C.c_int]),
            "fgl_copy_benchmark": (C.c_int, [P, C.c_size_t, C.c_int, C.POINTER(C.c_float)]),
        }
        for name, (res, args) in sigs.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib_handle = L
    return _lib_handle


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


# ---- host matrix builders (src/geometry.cpp:60-179) ----

def make_model_matrix(t, roty_deg, scale):
    out = np.zeros(16, np.float32)
    lib().fgl_make_model_matrix(float(t[0]), float(t[1]), float(t[2]), float(roty_deg), float(scale), _ptr(out))
    return out.reshape(4, 4)


def make_lookat(eye, center, up=(0.0, 1.0, 0.0)):
    out = np.zeros(16, np.float32)
    e, c, u = f32(eye), f32(center), f32(up)
    lib().fgl_make_lookat(_ptr(e), _ptr(c), _ptr(u), _ptr(out))
    return out.reshape(4, 4)


def make_perspective(fov, aspect, n, f):
    out = np.zeros(16, np.float32)
    lib().fgl_make_perspective(float(fov), float(aspect), float(n), float(f), _ptr(out))
    return out.reshape(4, 4)


def make_ortho(l, r, b, t, n, f):
    out = np.zeros(16, np.float32)
    lib().fgl_make_ortho(float(l), float(r), float(b), float(t), float(n), float(f), _ptr(out))
    return out.reshape(4, 4)


def mat_mul(a, b):
    out = np.zeros(16, np.float32)
    A, B = f32(a).reshape(16), f32(b).reshape(16)
    lib().fgl_mat_mul(_ptr(A), _ptr(B), _ptr(out))
    return out.reshape(4, 4)


def tga_write(path, img, rle=True):
    """Write B,G,R (h, w, 3) or grey (h, w) uint8 rows (row 0 = bottom) as TGA."""
    a = np.ascontiguousarray(img, dtype=np.uint8)
    bpp = 1 if a.ndim == 2 else a.shape[2]
    rc = lib().fgl_tga_write(os.fsencode(path), a.shape[1], a.shape[0], bpp, _ptr(a), int(rle))
    if rc != FGL_OK:
        raise FglError(f"fgl_tga_write({path}) failed: {rc}")


class Context:
    """One device context: resources, ForkerGL state and render targets."""

    def __init__(self, device=0):
        self._h = C.c_void_p()
        rc = lib().fgl_ctx_create(device, C.byref(self._h))
        if rc != FGL_OK:
            raise FglError(f"fgl_ctx_create failed: {rc}")
        self._keep = []

    def close(self):
        if self._h:
            lib().fgl_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc != FGL_OK:
            msg = lib().fgl_last_error(self._h)
            raise FglError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    # ---- resources ----
    def upload_texture(self, img_bgr, wrap=WRAP_NONE, filt=FILTER_NEAREST):
        a = np.ascontiguousarray(img_bgr, dtype=np.uint8)
        bpp = 1 if a.ndim == 2 else a.shape[2]
        out = C.c_int()
        self._chk(lib().fgl_upload_texture(self._h, a.shape[1], a.shape[0], bpp, _ptr(a), wrap, filt,
                                           C.byref(out)), "fgl_upload_texture")
        return out.value

    def upload_material(self, m):
        out = C.c_int()
        self._chk(lib().fgl_upload_material(self._h, C.byref(m), C.byref(out)), "fgl_upload_material")
        return out.value

    def upload_scene(self, sc):
        """Uploads every texture, material and model of a loaded Scene;
        returns [(model id, model matrix)] in the scene's model order."""
        tid = []
        for i in range(sc.info.n_textures):
            t = sc.texture(i)
            tid.append(self.upload_texture(t["bytes"], t["wrap"], t["filter"]))
        mid = []
        for i in range(sc.info.n_materials):
            mm = sc.material(i)
            for f in ("map_kd", "map_ks", "map_ke", "map_bump", "map_base", "map_rough", "map_metal", "map_ao",
                      "map_pbr_ke"):
                v = getattr(mm, f)
                if v >= 0:
                    setattr(mm, f, tid[v])
            mid.append(self.upload_material(mm))
        ids = []
        for i in range(sc.info.n_models):
            md, M = sc.model(i)
            md["face_material"] = np.array([mid[x] for x in md["face_material"]], np.int32)
            ids.append((self.upload_model(md), M))
        return ids

    def upload_model(self, model):
        """model: dict with pos (n,3), uv (n,2), nrm (n,3), tan (n,3) or None,
        vi/ti/ni (f,3), face_material (f,), support_pbr."""
        pos, uv, nrm = f32(model["pos"]), f32(model["uv"]), f32(model["nrm"])
        tan = None if model.get("tan") is None else f32(model["tan"])
        vi, ti, ni = i32(model["vi"]), i32(model["ti"]), i32(model["ni"])
        fm = i32(model["face_material"])
        d = ModelDesc()
        d.n_pos, d.n_uv, d.n_nrm, d.n_faces = len(pos.reshape(-1, 3)), len(uv.reshape(-1, 2)), \
            len(nrm.reshape(-1, 3)), len(vi.reshape(-1, 3))
        fp = C.POINTER(C.c_float)
        ip = C.POINTER(C.c_int32)
        d.pos, d.uv, d.nrm = pos.ctypes.data_as(fp), uv.ctypes.data_as(fp), nrm.ctypes.data_as(fp)
        d.tan = tan.ctypes.data_as(fp) if tan is not None else fp()
        d.vi, d.ti, d.ni = vi.ctypes.data_as(ip), ti.ctypes.data_as(ip), ni.ctypes.data_as(ip)
        d.face_material = fm.ctypes.data_as(ip)
        d.support_pbr = int(bool(model.get("support_pbr", False)))
        out = C.c_int()
        self._chk(lib().fgl_upload_model(self._h, C.byref(d), C.byref(out)), "fgl_upload_model")
        return out.value

    # ---- state ----
    def set_pass_tile_size(self, pass_, t):
        self._chk(lib().fgl_set_pass_tile_size(self._h, pass_, t), "fgl_set_pass_tile_size")

    def set_tile_size(self, t):
        self._chk(lib().fgl_set_tile_size(self._h, t), "fgl_set_tile_size")

    def init_targets(self, w, h):
        self._chk(lib().fgl_init_targets(self._h, w, h), "fgl_init_targets")

    def init_buffer(self, which, w, h):
        """ForkerGL::InitFrameBuffer / InitDepthBuffer / InitGeometryBuffers /
        InitShadowBuffer one at a time (which = TARGET_FRAME, _DEPTH, _GBUFFER, _SHADOW)."""
        self._chk(lib().fgl_init_buffer(self._h, which, w, h), "fgl_init_buffer")

    def init_shadow_buffer(self, w, h):
        self._chk(lib().fgl_init_shadow_buffer(self._h, w, h), "fgl_init_shadow_buffer")

    def clear_color(self, r, g, b):
        self._chk(lib().fgl_clear_color(self._h, r, g, b), "fgl_clear_color")

    def set_viewport(self, x, y, w, h, shadow=False):
        self._chk(lib().fgl_set_viewport(self._h, int(shadow), x, y, w, h), "fgl_set_viewport")

    def set_matrix(self, which, m):
        a = f32(m).reshape(16)
        self._chk(lib().fgl_set_matrix(self._h, which, _ptr(a)), "fgl_set_matrix")

    def set_eye(self, eye):
        a = f32(eye)
        self._chk(lib().fgl_set_eye(self._h, _ptr(a)), "fgl_set_eye")

    def set_lights(self, lights):
        arr = (Light * max(1, len(lights)))()
        for i, (typ, vec, col) in enumerate(lights):
            arr[i].type = typ
            arr[i].vec[:] = list(map(float, vec))
            arr[i].color[:] = list(map(float, col))
        self._chk(lib().fgl_set_lights(self._h, len(lights), arr), "fgl_set_lights")

    def set_shadow(self, on, cfg=None):
        self._chk(lib().fgl_set_shadow(self._h, int(on), C.byref(cfg) if cfg is not None else None),
                  "fgl_set_shadow")

    # ---- drawing ----
    def begin_pass(self, p):
        self._chk(lib().fgl_begin_pass(self._h, p), "fgl_begin_pass")

    def draw_model_normal(self, model_id, model_matrix, normal_matrix):
        a = f32(model_matrix).reshape(16)
        n = f32(normal_matrix).reshape(9)
        self._chk(lib().fgl_draw_model_normal(self._h, model_id, _ptr(a), _ptr(n)), "fgl_draw_model_normal")

    def draw_model(self, model_id, model_matrix):
        a = f32(model_matrix).reshape(16)
        self._chk(lib().fgl_draw_model(self._h, model_id, _ptr(a)), "fgl_draw_model")

    def end_pass(self):
        self._chk(lib().fgl_end_pass(self._h), "fgl_end_pass")

    def draw_screen_space_pixels(self):
        self._chk(lib().fgl_draw_screen_space_pixels(self._h), "fgl_draw_screen_space_pixels")

    def ssao(self, cfg):
        self._chk(lib().fgl_ssao(self._h, C.byref(cfg)), "fgl_ssao")

    def rng_seed(self, seed=5489):
        """Restarts the reference's random stream (std::mt19937(seed)) used by
        stochastic SSAO and the FGL_SHADOW_REF_* filters."""
        self._chk(lib().fgl_rng_seed(self._h, seed), "fgl_rng_seed")

    def rng_vectors(self, kind, n):
        """n RandomVectorInUnitSphere (kind 3) or RandomVectorInUnitDisk
        (kind 2) results from the stream, shape (n, kind)."""
        out = np.empty((n, kind), np.float32)
        self._chk(lib().fgl_rng_vectors(self._h, kind, n, out.ctypes.data_as(C.c_void_p)), "fgl_rng_vectors")
        return out

    def rng_draws(self):
        """32-bit draws consumed from the stream since its restart."""
        n = C.c_uint64()
        self._chk(lib().fgl_rng_draws(self._h, C.byref(n)), "fgl_rng_draws")
        return n.value

    def resolve(self, k=1):
        self._chk(lib().fgl_resolve(self._h, k), "fgl_resolve")

    def sync(self):
        self._chk(lib().fgl_sync(self._h), "fgl_sync")

    # ---- targets ----
    def target_bytes(self, which):
        n = C.c_size_t()
        self._chk(lib().fgl_target_bytes(self._h, which, C.byref(n)), "fgl_target_bytes")
        return n.value

    def read(self, which):
        n = self.target_bytes(which)
        dt = np.uint8 if which == TARGET_IMAGE else np.float32
        out = np.empty(n // np.dtype(dt).itemsize, dt)
        self._chk(lib().fgl_read_target(self._h, which, _ptr(out), n), "fgl_read_target")
        return out

    def write(self, which, arr):
        a = f32(arr).reshape(-1)
        self._chk(lib().fgl_write_target(self._h, which, _ptr(a), a.nbytes), "fgl_write_target")

    def check_recip(self):
        """(compared, mismatches, fast_path) of the kernels' reciprocal against
        the IEEE quotient over every non-NaN float."""
        a, b, f = C.c_uint64(), C.c_uint64(), C.c_uint64()
        self._chk(lib().fgl_check_recip(self._h, C.byref(a), C.byref(b), C.byref(f)), "fgl_check_recip")
        return a.value, b.value, f.value

    def sort_pairs(self, keys, vals, bits):
        """The binning's device radix sort (stable, by key) on copies of the
        arrays; returns the sorted (keys, vals)."""
        k = np.ascontiguousarray(keys, np.uint32).copy()
        v = np.ascontiguousarray(vals, np.uint32).copy()
        assert k.shape == v.shape and k.ndim == 1
        self._chk(lib().fgl_sort_pairs(self._h, _ptr(k), _ptr(v), k.size, bits), "fgl_sort_pairs")
        return k, v

    def enable_timing(self, on=True, only=None):
        """Kernel timing with HIP events; `only` restricts it to the kernels
        timed under that name (each event pair costs the stream a few us)."""
        self._chk(lib().fgl_time_only(self._h, only.encode() if only else None), "fgl_time_only")
        self._chk(lib().fgl_enable_timing(self._h, int(on)), "fgl_enable_timing")

    def read_timings(self, cap=4096):
        names = (C.c_char * 32 * cap)()
        ms = (C.c_float * cap)()
        n = C.c_int()
        self._chk(lib().fgl_read_timings(self._h, names, ms, cap, C.byref(n)), "fgl_read_timings")
        return [(bytes(names[i]).split(b"\0")[0].decode(), float(ms[i])) for i in range(min(n.value, cap))]

    def copy_benchmark(self, nbytes, iters=20):
        ms = C.c_float()
        self._chk(lib().fgl_copy_benchmark(self._h, nbytes, iters, C.byref(ms)), "fgl_copy_benchmark")
        return ms.value

    def render_scene(self, scene):
        self._chk(lib().fgl_render_scene(self._h, scene._h), "fgl_render_scene")


class Scene:
    """A .scene file loaded by the library's parsers (src/scene.cpp semantics)."""

    def __init__(self, path, wrap=None, filt=None):
        """wrap / filt: ForkerGL::TextureWrapping / TextureFiltering when the
        scene is loaded (default: the reference's NoWrap, Nearest); a
        `texture` key in the file overrides them."""
        self._h = C.c_void_p()
        err = C.create_string_buffer(1024)
        rc = lib().fgl_scene_load_with(os.fsencode(path), WRAP_NONE if wrap is None else wrap,
                                       FILTER_NEAREST if filt is None else filt, C.byref(self._h), err, 1024)
        if rc != FGL_OK:
            raise FglError(f"fgl_scene_load({path}) failed ({rc}): {err.value.decode()}")
        self.info = SceneInfo()
        lib().fgl_scene_info_get(self._h, C.byref(self.info))

    def __del__(self):
        try:
            if self._h:
                lib().fgl_scene_free(self._h)
        except Exception:
            pass

    def model(self, i):
        d = ModelDesc()
        mm = np.zeros(16, np.float32)
        if lib().fgl_scene_model(self._h, i, C.byref(d), _ptr(mm)) != FGL_OK:
            raise FglError(f"bad model index {i}")

        def arr(p, n, dt):
            if not p or n == 0:
                return None
            return np.ctypeslib.as_array(p, shape=(n,)).astype(dt, copy=True)

        out = {
            "pos": arr(d.pos, 3 * d.n_pos, np.float32).reshape(-1, 3),
            "uv": arr(d.uv, 2 * d.n_uv, np.float32).reshape(-1, 2),
            "nrm": arr(d.nrm, 3 * d.n_nrm, np.float32).reshape(-1, 3),
            "tan": None if not d.tan else arr(d.tan, 3 * d.n_pos, np.float32).reshape(-1, 3),
            "vi": arr(d.vi, 3 * d.n_faces, np.int32).reshape(-1, 3),
            "ti": arr(d.ti, 3 * d.n_faces, np.int32).reshape(-1, 3),
            "ni": arr(d.ni, 3 * d.n_faces, np.int32).reshape(-1, 3),
            "face_material": arr(d.face_material, d.n_faces, np.int32),
            "support_pbr": bool(d.support_pbr),
        }
        return out, mm.reshape(4, 4)

    def material(self, i):
        m = Material()
        if lib().fgl_scene_material(self._h, i, C.byref(m)) != FGL_OK:
            raise FglError(f"bad material index {i}")
        return m

    def texture(self, i):
        w, h, bpp, wrap, filt = C.c_int(), C.c_int(), C.c_int(), C.c_int(), C.c_int()
        p = C.c_void_p()
        if lib().fgl_scene_texture(self._h, i, C.byref(w), C.byref(h), C.byref(bpp), C.byref(wrap),
                                   C.byref(filt), C.byref(p)) != FGL_OK:
            raise FglError(f"bad texture index {i}")
        n = w.value * h.value * bpp.value
        data = np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_uint8)), shape=(n,)).copy()
        shape = (h.value, w.value) if bpp.value == 1 else (h.value, w.value, bpp.value)
        return {"bytes": data.reshape(shape), "wrap": wrap.value, "filter": filt.value, "bpp": bpp.value}
