"""Oracle restatement of the reference's file loaders (TEST INFRASTRUCTURE).

Independent of the product's C++ parsers: parses .scene / .obj / .mtl / .tga
the way src/scene.cpp:19-196, src/model.cpp:16-476 and
src/tgaimage.cpp:43-168 do, and returns the same arrays the product exposes
through fgl_scene_model / fgl_scene_material / fgl_scene_texture, so tests can
compare the two loaders array for array.

Float parsing goes through float64 then float32 (numpy); for the decimal
strings in the assets (<= 9 significant digits) this equals the correctly
rounded float32 that istream >> float produces.
"""
import os
import struct

import numpy as np

F32 = np.float32


def _ltrim(s):
    return s.lstrip(" \n\r\t\f\v")


# ---------------- TGA (src/tgaimage.cpp:43-168, 350-365) ----------------

def tga_read(path):
    with open(path, "rb") as fh:
        data = fh.read()
    if len(data) < 18:
        raise IOError("short header")
    (idlen, cmtype, dtype, cmo, cml, cmd, xo, yo, w, h, bits, desc) = struct.unpack("<BBBHHBHHHHBB", data[:18])
    bpp = bits >> 3
    if w <= 0 or h <= 0 or bpp not in (1, 3, 4):
        raise IOError("bad bpp (or width/height) value")
    n = w * h * bpp
    pos = 18  # the reference ignores the image ID field
    if dtype in (2, 3):
        raw = np.frombuffer(data[pos:pos + n], np.uint8)
        if raw.size != n:
            raise IOError("short data")
    elif dtype in (10, 11):
        out = bytearray()
        npix = w * h
        cur = 0
        while cur < npix:
            ch = data[pos]
            pos += 1
            if ch < 128:
                cnt = ch + 1
                chunk = data[pos:pos + cnt * bpp]
                pos += cnt * bpp
                out += chunk
            else:
                cnt = ch - 127
                px = data[pos:pos + bpp]
                pos += bpp
                out += px * cnt
            cur += cnt
            if cur > npix:
                raise IOError("too many pixels read")
        raw = np.frombuffer(bytes(out), np.uint8)
    else:
        raise IOError(f"unknown file format {dtype}")
    img = raw.reshape(h, w, bpp).copy()
    if not (desc & 0x20):
        img = img[::-1]
    if desc & 0x10:
        img = img[:, ::-1]
    return np.ascontiguousarray(img)


# ---------------- math helpers in float32 ----------------

def tga_rle_encode(data, bpp):
    """Packet stream of TGAImage::unloadRleData (src/tgaimage.cpp:250-300) for
    `data`, the pixel bytes in file order: the reference's greedy loop, step
    for step -- grow the packet while the next pixel compares as the first
    pair did, a literal packet giving back its last pixel when it meets an
    equal pair, at most 128 pixels per packet."""
    data = bytes(data)
    npix = len(data) // bpp
    out = bytearray()
    cur = 0
    while cur < npix:
        start = cur * bpp
        byte = cur * bpp
        run = 1
        raw = True
        while cur + run < npix and run < 128:
            eq = data[byte:byte + bpp] == data[byte + bpp:byte + 2 * bpp]
            byte += bpp
            if run == 1:
                raw = not eq
            if raw and eq:
                run -= 1
                break
            if not raw and not eq:
                break
            run += 1
        cur += run
        out.append(run - 1 if raw else run + 127)
        out += data[start:start + (run * bpp if raw else bpp)]
    return bytes(out)


def _norm3(v):
    v = np.asarray(v, F32)
    ls = F32(F32(v[0] * v[0]) + F32(v[1] * v[1]))
    ls = F32(ls + F32(v[2] * v[2]))
    inv = F32(F32(1.0) / np.sqrt(ls, dtype=F32))
    return np.array([v[0] * F32(1.0) * inv, v[1] * F32(1.0) * inv, v[2] * F32(1.0) * inv], F32)


def _len3(v):
    ls = F32(F32(v[0] * v[0]) + F32(v[1] * v[1]))
    return np.sqrt(F32(ls + F32(v[2] * v[2])), dtype=F32)


# ---------------- OBJ / MTL ----------------

def _default_material():
    return {"ka": [0.0] * 3, "kd": [0.0] * 3, "ks": [0.0] * 3, "ke": [0.0] * 3,
            "map_kd": -1, "map_ks": -1, "map_ke": -1, "map_bump": -1,
            "pbr_ke": [0.0] * 3, "albedo": [1.0] * 3, "roughness": 0.0, "metalness": 0.0,
            "map_base": -1, "map_rough": -1, "map_metal": -1, "map_ao": -1, "map_pbr_ke": -1}


class Loader:
    def __init__(self):
        self.textures = []      # dicts: bytes, wrap, filter
        self._tex_cache = {}
        self.materials = []
        self.wrap = 0
        self.filter = 0

    def texture(self, path):
        key = (path, self.wrap, self.filter)
        if key in self._tex_cache:
            return self._tex_cache[key]
        try:
            img = tga_read(path)
        except (IOError, OSError):
            self._tex_cache[key] = -1
            return -1
        img = np.ascontiguousarray(img[::-1])  # flipTexCoordY (src/model.cpp:423)
        if img.shape[2] == 1:
            img = img[:, :, 0]
        self.textures.append({"bytes": img, "wrap": self.wrap, "filter": self.filter})
        self._tex_cache[key] = len(self.textures) - 1
        return self._tex_cache[key]

    def mtl(self, directory, fname, mats):
        try:
            fh = open(directory + fname)
        except OSError:
            return
        name = None
        for line in fh.read().split("\n"):
            line = _ltrim(line)
            tok = line.split()
            if line.startswith("newmtl "):
                name = tok[1] if len(tok) > 1 else ""
                mats[name] = _default_material()
                continue
            if name not in mats:
                continue
            m = mats[name]

            def vec():
                v = [0.0, 0.0, 0.0]
                for i in range(3):
                    try:
                        v[i] = float(F32(float(tok[1 + i])))
                    except (IndexError, ValueError):
                        break
                return v

            def tex():
                return self.texture(directory + tok[1]) if len(tok) > 1 else -1
            if line.startswith("Ka "):
                m["ka"] = vec()
            elif line.startswith("Kd "):
                m["kd"] = vec(); m["albedo"] = list(m["kd"])
            elif line.startswith("Ks "):
                m["ks"] = vec()
            elif line.startswith("Ke "):
                m["ke"] = vec(); m["pbr_ke"] = list(m["ke"])
            elif line.startswith("Pr "):
                m["roughness"] = float(F32(float(tok[1])))
            elif line.startswith("Pm "):
                m["metalness"] = float(F32(float(tok[1])))
            elif line.startswith("map_Kd "):
                t = tex()
                if t >= 0:
                    m["map_kd"] = m["map_base"] = t
            elif line.startswith("map_Ks "):
                t = tex()
                if t >= 0:
                    m["map_ks"] = t
            elif line.startswith("map_Ke "):
                t = tex()
                if t >= 0:
                    m["map_ke"] = m["map_pbr_ke"] = t
            elif line.startswith("map_Bump ") or line.startswith("norm "):
                t = tex()
                if t >= 0:
                    m["map_bump"] = t
            elif line.startswith("map_Ao "):
                t = tex()
                if t >= 0:
                    m["map_ao"] = t
            elif line.startswith("map_Pr "):
                t = tex()
                if t >= 0:
                    m["map_rough"] = t
            elif line.startswith("map_Pm "):
                t = tex()
                if t >= 0:
                    m["map_metal"] = t

    def obj(self, path, normalized, gen_tangent):
        pos, uv, nrm = [], [], []
        meshes = {}
        mats = {}
        mesh = None
        directory = path[:path.rfind("/") + 1] if "/" in path else ""
        with open(path) as fh:
            lines = fh.read().split("\n")
        for line in lines:
            line = _ltrim(line)
            tok = line.split()

            def floats(n):
                v = [0.0] * n
                for i in range(n):
                    try:
                        v[i] = float(F32(float(tok[1 + i])))
                    except (IndexError, ValueError):
                        break
                return v
            if line.startswith("mtllib "):
                self.mtl(directory, tok[1], mats)
            elif line.startswith("v "):
                pos.append(floats(3))
            elif line.startswith("vt "):
                uv.append(floats(2))
            elif line.startswith("vn "):
                nrm.append(floats(3))
            elif line.startswith("g "):
                mesh = tok[1] if len(tok) > 1 else ""
                meshes[mesh] = {"f": [], "mat": None}
            elif line.startswith("usemtl "):
                meshes[mesh]["mat"] = tok[1]
            elif line.startswith("f "):
                verts = []
                for t in tok[1:]:
                    parts = t.split("/")
                    try:
                        verts.append(tuple(int(p) - 1 for p in parts[:3]))
                    except ValueError:
                        break
                    if len(parts) < 3:
                        break
                for i in range(1, len(verts) - 1):   # fan (src/model.cpp:234-250)
                    meshes[mesh]["f"].append((verts[0], verts[i], verts[i + 1]))
        # materials appended in std::map (sorted name) order, model-local ids
        mat_id = {}
        for name in sorted(mats, key=lambda s: s.encode()):
            self.materials.append(mats[name])
            mat_id[name] = len(self.materials) - 1
        vi, ti, ni, fm = [], [], [], []
        support_pbr = False
        for name in sorted(meshes, key=lambda s: s.encode()):   # std::map order
            mm = meshes[name]
            mid = mat_id[mm["mat"]]
            for f in mm["f"]:
                vi.append([f[0][0], f[1][0], f[2][0]])
                ti.append([f[0][1], f[1][1], f[2][1]])
                ni.append([f[0][2], f[1][2], f[2][2]])
                fm.append(mid)
            m = self.materials[mid]
            support_pbr = m["map_metal"] >= 0 or m["map_rough"] >= 0
        P = np.array(pos, F32).reshape(-1, 3)
        T = np.array(uv, F32).reshape(-1, 2)
        N = np.array(nrm, F32).reshape(-1, 3)
        VI, TI, NI = (np.array(x, np.int32).reshape(-1, 3) for x in (vi, ti, ni))
        tan = generate_tangents(P, T, VI, TI) if gen_tangent else None
        if normalized:
            P = normalize_positions(P)
        return {"pos": P, "uv": T, "nrm": N, "tan": tan, "vi": VI, "ti": TI, "ni": NI,
                "face_material": np.array(fm, np.int32), "support_pbr": support_pbr}


def generate_tangents(P, T, VI, TI):
    """src/model.cpp:430-476, in float32 scalar arithmetic."""
    acc = np.zeros((len(P), 3), F32)
    for f in range(len(VI)):
        p0, p1, p2 = VI[f]
        e1 = P[p1] - P[p0]
        e2 = P[p2] - P[p0]
        t0, t1, t2 = T[TI[f][0]], T[TI[f][1]], T[TI[f][2]]
        d1 = t1 - t0
        d2 = t2 - t0
        det = F32(F32(d1[0] * d2[1]) - F32(d2[0] * d1[1]))
        if det == F32(0):
            continue
        inv = F32(F32(1.0) / det)
        v = np.array([F32(d2[1] * e1[0]) - F32(d1[1] * e2[0]), F32(d2[1] * e1[1]) - F32(d1[1] * e2[1]),
                      F32(d2[1] * e1[2]) - F32(d1[1] * e2[2])], F32)
        Tv = _norm3(v * inv)
        acc[p0] += Tv
        acc[p1] += Tv
        acc[p2] += Tv
    out = np.empty_like(acc)
    for i in range(len(acc)):
        out[i] = np.array([1, 0, 0], F32) if _len3(acc[i]) == F32(0) else _norm3(acc[i])
    return out


def normalize_positions(P):
    """src/model.cpp:107-138 (maxima start at FLT_MIN; translation not scaled)."""
    fmax, fmin = np.finfo(F32).max, np.finfo(F32).tiny
    mn = [fmax] * 3
    mx = [fmin] * 3
    for v in P:
        for k in range(3):
            mn[k] = v[k] if v[k] < mn[k] else mn[k]
            mx[k] = v[k] if mx[k] < v[k] else mx[k]
    ext = [F32(mx[k] - mn[k]) for k in range(3)]
    # Max(a, b) = (a < b) ? b : a
    inner = ext[2] if ext[1] < ext[2] else ext[1]
    big = inner if ext[0] < inner else ext[0]
    sc = F32(F32(2.0) / big)
    t = [F32(F32(-(mx[k] + mn[k])) / ext[k]) for k in range(3)]
    out = np.empty_like(P)
    for i, v in enumerate(P):
        for k in range(3):
            # row k of M: sc at k, t[k] at column 3: ((0 + sc*v_k) + 0*..) + t_k*1
            r = F32(0.0)
            for j in range(3):
                r = F32(r + F32((sc if j == k else F32(0.0)) * v[j]))
            r = F32(r + F32(t[k] * F32(1.0)))
            out[i, k] = r
    return out
