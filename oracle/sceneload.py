"""Oracle restatement of the .scene parser (src/scene.cpp:19-196) on top of
oracle.objload (TEST INFRASTRUCTURE).

Returns the models (arrays + model matrix, one entry per "model" line), the
scene-global material and texture tables and the reference keys
(screen, ssaa, ssao, shadow, mode, lights, camera).  Extension keys the
product reads (shadowmap, shadowfilter, ssaokernel, aoblur, multilight) are
sample-set generators and are taken from the product's scene info by the
tests; "texture" changes the wrap/filter of later textures exactly as
ForkerGL::TextureWrapMode / TextureFilterMode before a load would.
"""
import numpy as np

from . import make_model_matrix
from .objload import Loader, _ltrim

WRAP = {"repeat": 1, "mirror": 2, "clamp": 3}


def load_scene(path):
    L = Loader()
    info = {"width": 1280, "height": 800, "ssaa_on": 0, "ssaa_k": 2, "ssao_on": 0, "shadow_on": 1,
            "deferred": 0, "camera_ortho": 0, "eye": [0.0] * 3, "look_at": [0.0] * 3,
            "point": None, "dir": None}
    models = []
    cache = {}
    with open(path) as fh:
        lines = fh.read().split("\n")
    for line in lines:
        line = _ltrim(line)
        tok = line.split()
        if line.startswith("#"):
            continue
        if line.startswith("mode "):
            info["deferred"] = int(len(tok) > 1 and tok[1] == "deferred")
        elif line.startswith("screen "):
            info["width"], info["height"] = int(tok[1]), int(tok[2])
        elif line.startswith("ssaa "):
            info["ssaa_on"] = int(tok[1] == "on")
            if len(tok) > 2:
                info["ssaa_k"] = int(tok[2])
        elif line.startswith("ssao "):
            info["ssao_on"] = int(tok[1] == "on")
        elif line.startswith("shadow "):
            info["shadow_on"] = int(tok[1] == "on")
        elif line.startswith("light "):
            vals = [float(np.float32(float(x))) for x in tok[2:8]]
            if tok[1] == "point":
                info["point"] = vals
            elif tok[1] == "dir":
                info["dir"] = vals
        elif line.startswith("camera "):
            if tok[1] not in ("persp", "ortho"):
                continue
            info["camera_ortho"] = int(tok[1] == "ortho")
            v = [float(np.float32(float(x))) for x in tok[2:8]]
            info["eye"], info["look_at"] = v[:3], v[3:6]
        elif line.startswith("model "):
            f, norm, gen = tok[1], tok[2] == "true", tok[3] == "true"
            p = [float(np.float32(float(x))) for x in tok[4:7]]
            rot, scale = float(np.float32(float(tok[7]))), float(np.float32(float(tok[8])))
            key = (f, norm, gen, L.wrap, L.filter)
            if key not in cache:
                cache[key] = L.obj(f, norm, gen)
            models.append((cache[key], make_model_matrix(p, rot, scale)))
        elif line.startswith("texture "):
            L.wrap = WRAP.get(tok[1], 0)
            L.filter = 1 if tok[2] == "linear" else 0
    return {"models": models, "materials": L.materials, "textures": L.textures, "info": info}
