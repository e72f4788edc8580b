"""Seeded synthetic scenes for the benchmark configurations.

Every scene is written as files the reference renderer can also read: a
.scene (with extension keys the reference parser skips), .obj meshes with
"g"/"usemtl" groups, .mtl materials and uncompressed TGA textures.  Floats
are written with 9 significant digits, so parsing recovers the float32 values
exactly.  The same seed gives the same bytes on every machine.
"""
import math
import os

import numpy as np

from . import tga_write

GOLDEN = (1 + 5 ** 0.5) / 2


# ---------------- geometry ----------------

def uv_sphere(n_lon=64, n_lat=32, radius=1.0):
    """UV sphere with duplicated seam: (n_lat+1)*(n_lon+1) vertices, normals =
    unit position, uv = (j/n_lon, i/n_lat); pole bands emit one triangle per
    segment (no degenerate triangles)."""
    i = np.arange(n_lat + 1)[:, None]
    j = np.arange(n_lon + 1)[None, :]
    theta = np.pi * i / n_lat            # 0 at north pole
    phi = 2 * np.pi * j / n_lon
    x = np.sin(theta) * np.cos(phi)
    y = np.cos(theta) * np.ones_like(phi)
    z = -np.sin(theta) * np.sin(phi)
    nrm = np.stack([x, y, z], -1).reshape(-1, 3)
    pos = nrm * radius
    shape = (n_lat + 1, n_lon + 1)
    uv = np.stack([np.broadcast_to(j / n_lon, shape), np.broadcast_to(1 - i / n_lat, shape)], -1).reshape(-1, 2)
    faces = []
    W = n_lon + 1
    for a in range(n_lat):
        for b in range(n_lon):
            v00, v01 = a * W + b, a * W + b + 1
            v10, v11 = (a + 1) * W + b, (a + 1) * W + b + 1
            if a != 0:
                faces.append((v00, v10, v01))
            if a != n_lat - 1:
                faces.append((v01, v10, v11))
    return pos, nrm, uv, np.array(faces, np.int64)


def icosphere(subdiv):
    t = GOLDEN
    v = np.array([(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t),
                  (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)], np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
                  (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
                  (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)], np.int64)
    for _ in range(subdiv):
        n = len(v)
        e = f[:, [0, 1, 1, 2, 2, 0]].reshape(-1, 2)
        lo, hi = e.min(1), e.max(1)
        key = lo * n + hi
        uniq, inv = np.unique(key, return_inverse=True)
        a, b = uniq // n, uniq % n
        mid = v[a] + v[b]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        v = np.concatenate([v, mid])
        m = (inv.reshape(-1, 3) + n)
        m01, m12, m20 = m[:, 0], m[:, 1], m[:, 2]
        c0, c1, c2 = f[:, 0], f[:, 1], f[:, 2]
        f = np.concatenate([np.stack([c0, m01, m20], 1), np.stack([c1, m12, m01], 1),
                            np.stack([c2, m20, m12], 1), np.stack([m01, m12, m20], 1)])
    return v, f


def _hash3(ix, iy, iz, seed):
    h = (ix.astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15) ^ iy.astype(np.uint64) * np.uint64(0xC2B2AE3D27D4EB4F)
         ^ iz.astype(np.uint64) * np.uint64(0x165667B19E3779F9) ^ np.uint64(seed * 0x27D4EB2F165667C5 & (2**64 - 1)))
    h ^= h >> np.uint64(33)
    h *= np.uint64(0xFF51AFD7ED558CCD)
    h ^= h >> np.uint64(33)
    h *= np.uint64(0xC4CEB9FE1A85EC53)
    h ^= h >> np.uint64(33)
    return (h >> np.uint64(11)).astype(np.float64) / float(1 << 53)


def value_noise(p, seed, octaves=3, freq=2.0, lacunarity=2.0, gain=0.5):
    """Sum of `octaves` trilinear (smoothstep) value-noise layers in [-1, 1]."""
    total = np.zeros(len(p))
    amp, norm = 1.0, 0.0
    with np.errstate(over="ignore"):
        for o in range(octaves):
            q = p * freq + 17.0 * (o + 1)
            i0 = np.floor(q).astype(np.int64)
            fr = q - i0
            s = fr * fr * (3 - 2 * fr)
            acc = np.zeros(len(p))
            for dx in (0, 1):
                for dy in (0, 1):
                    for dz in (0, 1):
                        w = (s[:, 0] if dx else 1 - s[:, 0]) * (s[:, 1] if dy else 1 - s[:, 1]) * \
                            (s[:, 2] if dz else 1 - s[:, 2])
                        acc += w * _hash3(i0[:, 0] + dx, i0[:, 1] + dy, i0[:, 2] + dz, seed + o)
            total += amp * (acc * 2 - 1)
            norm += amp
            amp *= gain
            freq *= lacunarity
    return total / norm


def scatter_add(index, values, n):
    """Sum of the rows of `values` (m, c) into n rows by `index` (m,)."""
    return np.stack([np.bincount(index, weights=values[:, c], minlength=n) for c in range(values.shape[1])], 1)


def vertex_normals(v, f):
    fn = np.cross(v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 0]])
    n = sum(scatter_add(f[:, k], fn, len(v)) for k in range(3))
    ln = np.linalg.norm(n, axis=1, keepdims=True)
    ln[ln == 0] = 1
    return n / ln


_ICO = {}


def displaced_icosphere(subdiv, seed, amplitude=0.1, radius=1.0):
    if subdiv not in _ICO:
        _ICO[subdiv] = icosphere(subdiv)
    v, f = _ICO[subdiv]
    d = value_noise(v, seed)
    v = v * (radius * (1 + amplitude * d))[:, None]
    n = vertex_normals(v, f)
    u = np.arctan2(v[:, 2], v[:, 0]) / (2 * np.pi) + 0.5
    w = np.arcsin(np.clip(v[:, 1] / np.linalg.norm(v, axis=1), -1, 1)) / np.pi + 0.5
    uv = np.stack([u, w], 1)
    return v, n, uv, f


# ---------------- writers ----------------

def _fmt(a):
    return "\n".join(" ".join(f"{x:.9g}" for x in row) for row in np.asarray(a, np.float32))


def write_obj(path, groups, mtllib):
    """groups: list of (name, material, pos, nrm, uv, faces) sharing nothing;
    indices are offset per group."""
    lines = [f"mtllib {mtllib}"]
    off_v = 0
    body = []
    for name, mat, pos, nrm, uv, faces in groups:
        lines.append("\n".join("v " + s for s in _fmt(pos).split("\n")))
        lines.append("\n".join("vt " + s for s in _fmt(uv).split("\n")))
        lines.append("\n".join("vn " + s for s in _fmt(nrm).split("\n")))
        fi = np.asarray(faces, np.int64) + off_v + 1
        fl = [f"f {a}/{a}/{a} {b}/{b}/{b} {c}/{c}/{c}" for a, b, c in fi]
        body.append(f"g {name}\nusemtl {mat}\n" + "\n".join(fl))
        off_v += len(pos)
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n" + "\n".join(body) + "\n")


def write_mtl(path, mats):
    out = []
    for name, kv in mats:
        out.append(f"newmtl {name}")
        for k, v in kv:
            out.append(f"    {k} {v}")
    with open(path, "w") as fh:
        fh.write("\n".join(out) + "\n")


def _tga_rows(img):
    """numpy image given top row first (h, w, 3 RGB) -> B,G,R rows bottom-first."""
    a = np.asarray(img, np.uint8)
    if a.ndim == 3:
        a = a[..., ::-1]
    return np.ascontiguousarray(a[::-1])


def write_tga(path, img):
    tga_write(path, _tga_rows(img), rle=False)


def plane_group(scale_uv=1.0):
    pos = np.array([(-1, 0, 1), (1, 0, 1), (-1, 0, -1), (1, 0, -1)], np.float32)
    nrm = np.array([(0, 1, 0)] * 4, np.float32)
    uv = np.array([(0, 0), (1, 0), (0, 1), (1, 1)], np.float32) * scale_uv
    faces = np.array([(0, 1, 2), (1, 3, 2)])
    return pos, nrm, uv, faces


def pbr_textures(d, seed, size):
    rng = np.random.default_rng(seed)
    write_tga(os.path.join(d, "albedo.tga"), rng.integers(0, 256, (size, size, 3), dtype=np.uint8))
    rough = np.round(255 * rng.uniform(0.05, 1.0, (size, size))).astype(np.uint8)
    write_tga(os.path.join(d, "roughness.tga"), rough)
    metal = (rng.random((size, size)) < 0.5).astype(np.uint8) * 255
    write_tga(os.path.join(d, "metalness.tga"), metal)
    nxy = rng.normal(0, 0.2, (size, size, 2))
    n = np.concatenate([nxy, np.ones((size, size, 1))], -1)
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    write_tga(os.path.join(d, "normal.tga"), np.round((n + 1) / 2 * 255).astype(np.uint8))


# Seed of the 64-sample SSAO kernel and 4 x 4 noise of the deferred configs:
# the heavy 1080p configuration uses the 1280 x 800 configuration's SSAO.
SSAO_KERNEL_SEED = 2

PBR_MTL = [("Ka", "1 1 1"), ("Kd", "1 1 1"), ("Ks", "0.5 0.5 0.5"), ("map_Kd", "albedo.tga"),
           ("map_Pr", "roughness.tga"), ("map_Pm", "metalness.tga"), ("map_Bump", "normal.tga")]


# ---------------- configurations ----------------

def config1(d, seed=1):
    """Forward Blinn-Phong 640x400, hard shadow, plane + 64x32 UV sphere with a
    512^2 noise texture (repeat, bilinear)."""
    os.makedirs(d, exist_ok=True)
    rng = np.random.default_rng(seed)
    write_obj(os.path.join(d, "plane.obj"), [("Plane", "Plane", *plane_group())], "plane.mtl")
    write_mtl(os.path.join(d, "plane.mtl"), [("Plane", [("Ka", "0.5 0.5 0.5"), ("Kd", "0.5 0.5 0.5"),
                                                        ("Ks", "0.2 0.2 0.2")])])
    pos, nrm, uv, f = uv_sphere(64, 32, 1.0)
    write_obj(os.path.join(d, "sphere.obj"), [("Sphere", "Noise", pos, nrm, uv, f)], "sphere.mtl")
    write_mtl(os.path.join(d, "sphere.mtl"), [("Noise", [("Ka", "1 1 1"), ("Kd", "1 1 1"), ("Ks", "0.5 0.5 0.5"),
                                                         ("map_Kd", "noise.tga")])])
    write_tga(os.path.join(d, "noise.tga"), rng.integers(0, 256, (512, 512, 3), dtype=np.uint8))
    scene = f"""# config 1: forward Blinn-Phong, hard shadow
mode forward
screen 640 400
ssaa off 2
ssao off
shadow on
shadowfilter hard
light point 2 5 5 1 1 1
camera persp -1 1 1 0 0 -1
model {d}/plane.obj false false 0 -1 -1 0 3
texture repeat linear
model {d}/sphere.obj false false 0 0 -3 0 1
"""
    p = os.path.join(d, "c1.scene")
    with open(p, "w") as fh:
        fh.write(scene)
    return p


def _pbr_scene_assets(d, seed, subdiv, tex_size, radius=1.0, ground_scale=4.0):
    os.makedirs(d, exist_ok=True)
    v, n, uv, f = displaced_icosphere(subdiv, seed, 0.1, radius)
    write_obj(os.path.join(d, "sphere.obj"), [("Sphere", "Pbr", v, n, uv, f)], "pbr.mtl")
    write_obj(os.path.join(d, "ground.obj"), [("Ground", "Pbr", *plane_group(4.0))], "pbr.mtl")
    write_mtl(os.path.join(d, "pbr.mtl"), [("Pbr", PBR_MTL)])
    pbr_textures(d, seed + 1000, tex_size)


def config2(d, seed=2, subdiv=7, tex_size=2048, width=1280, height=800, camera=None):
    """Deferred PBR: displaced icosphere on a ground plane, 2048^2 PBR maps,
    2048^2 shadow map with 5x5 PCF, 64-sample kernel SSAO r=0.5, sigma=2 blur."""
    _pbr_scene_assets(d, seed, subdiv, tex_size)
    eye, at = camera if camera is not None else ((0.0, 1.2, 3.2), (0.0, -0.2, 0.0))
    scene = f"""# config 2: deferred PBR + PCF 5x5 + SSAO
mode deferred
screen {width} {height}
ssaa off 2
ssao on
shadow on
shadowmap 2048
shadowfilter grid 5
ssaokernel 64 0.5 0.0005 0.01 {SSAO_KERNEL_SEED}
aoblur gaussian 2
light point 2 5 5 3 3 3
camera persp {eye[0]} {eye[1]} {eye[2]} {at[0]} {at[1]} {at[2]}
texture repeat linear
model {d}/ground.obj true true 0 -1.1 0 0 3
model {d}/sphere.obj false true 0 0 0 0 1
"""
    p = os.path.join(d, "c2.scene")
    with open(p, "w") as fh:
        fh.write(scene)
    return p


def config3(d, seed=3, subdiv=6, tex_size=2048, width=1920, height=1080, ssaa=2, grid=8, name="c3"):
    """Heavy deferred PBR at 1920x1080 with SSAA 2x, PCSS 16/32 area 0.05, kernel
    SSAO as config 2, an 8x8 grid of displaced icospheres with seeded scale and
    rotation, one directional and one point light."""
    _pbr_scene_assets(d, seed, subdiv, tex_size, radius=0.3)
    rng = np.random.default_rng(seed)
    lines = []
    spacing = 0.8
    for gz in range(grid):
        for gx in range(grid):
            s = rng.uniform(0.5, 1.5)
            r = rng.uniform(0, 360)
            x = (gx - (grid - 1) / 2) * spacing
            z = (gz - (grid - 1) / 2) * spacing
            lines.append(f"model {d}/sphere.obj false true {x:.6g} 0 {z:.6g} {r:.6g} {s:.6g}")
    scene = f"""# config 3: heavy deferred PBR, SSAA 2x, PCSS, SSAO, 64 instances
mode deferred
screen {width} {height}
ssaa {'on' if ssaa > 1 else 'off'} {max(ssaa, 1)}
ssao on
shadow on
shadowmap 2048
shadowfilter pcss 16 32 0.05
ssaokernel 64 0.5 0.0005 0.01 {SSAO_KERNEL_SEED}
aoblur gaussian 2
multilight on
light point 2 5 5 2.5 2.5 2.5
light dir -0.4 -1 -0.3 0.6 0.6 0.7
camera persp 0 3.6 5.4 0 -0.3 0
texture repeat linear
model {d}/ground.obj false true 0 -0.45 0 0 4
""" + "\n".join(lines) + "\n"
    p = os.path.join(d, f"{name}.scene")
    with open(p, "w") as fh:
        fh.write(scene)
    return p


def config4(d, seed=3):
    """Config 3 without SSAA (1920x1080 internal): the neural-AO comparison frame."""
    return config3(d, seed=seed, ssaa=1, name="c4")


def tiny_scene(d, seed=0, width=64, height=48, mode="deferred", shadow="hard", ssao=False,
               subdiv=1, tex=16, ssaa=1, pbr=True, extra="", texture="repeat linear", shadowmap=128, uv_scale=1.0):
    """Small scene for parity tests: a displaced icosphere and a ground plane.
    texture: the `texture WRAP FILTER` key, or None for the reference's
    load-time defaults; shadowmap: the `shadowmap S` key, or None for a
    frame-sized shadow map (the reference's); uv_scale multiplies the
    sphere's and the ground's texture coordinates (values outside [0, 1]
    exercise the wrap modes)."""
    os.makedirs(d, exist_ok=True)
    v, n, uv, f = displaced_icosphere(subdiv, seed, 0.1, 0.8)
    rng = np.random.default_rng(seed)
    mat = "Pbr" if pbr else "Phong"
    gp, gn, guv, gf = plane_group(2.0)
    if uv_scale != 1.0:
        uv = (uv - 0.25) * uv_scale
        guv = (guv - 0.5) * uv_scale
    write_obj(os.path.join(d, "sphere.obj"), [("Sphere", mat, v, n, uv, f)], "m.mtl")
    write_obj(os.path.join(d, "ground.obj"), [("Ground", mat, gp, gn, guv, gf)], "m.mtl")
    phong = [("Ka", "0.2 0.2 0.2"), ("Kd", "0.7 0.6 0.5"), ("Ks", "0.4 0.4 0.4"), ("map_Kd", "albedo.tga"),
             ("map_Ks", "roughness.tga"), ("map_Bump", "normal.tga")]
    write_mtl(os.path.join(d, "m.mtl"), [("Pbr", PBR_MTL), ("Phong", phong)])
    pbr_textures(d, seed + 7, tex)
    filt = {"hard": "hard", "grid": "grid 3", "pcss": "pcss 8 16 0.5", "pcf": "pcf 16 0.01"}[shadow]
    scene = f"""mode {mode}
screen {width} {height}
ssaa {'on' if ssaa > 1 else 'off'} {max(ssaa, 1)}
ssao {'on' if ssao else 'off'}
shadow on
{"" if shadowmap is None else f"shadowmap {shadowmap}"}
shadowfilter {filt}
ssaokernel 16 0.5 0.0005 0.01 {seed}
aoblur gaussian 2
light point 2 5 5 2 2 2
camera persp 0.3 1.0 2.6 0 -0.2 0
{"" if texture is None else f"texture {texture}"}
model {d}/ground.obj false true 0 -0.9 0 0 2
model {d}/sphere.obj false true 0 0 0 {seed * 37 % 360} 1
{extra}
"""
    p = os.path.join(d, "tiny.scene")
    with open(p, "w") as fh:
        fh.write(scene)
    return p
