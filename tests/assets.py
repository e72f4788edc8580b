"""The reference's sample assets (obj/<name>/), kept as byte-for-byte
fixtures under tests/golden/assets, unpacked into a directory so the
product's .scene/.obj/.mtl/.tga parsers open them as files, and .scene
files that place them as the reference's scenes/test.scene does."""
import os

import numpy as np

FIXTURES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "assets")

# model line arguments (normalise, generate tangents, x y z, rotate_y, scale):
# the placements of scenes/test.scene:28-63 (its commented-out lines
# included); brickwall is not in that file and stands where mary does.
PLACEMENT = {
    "plane": ("plane.obj", "false false 0 -1 -1 0 3"),
    "diablo_pose": ("diablo_pose.obj", "true true 0 0 0 -10 1"),
    "mary": ("mary.obj", "true true 0.05 0 -1 -10 1"),
    "great_sword": ("great_sword.obj", "true true 0 0.3 -1 240 1"),
    "african_head": ("head.obj", "true true 0 0 0 -1 0.5"),
    "african_head/eye": ("eye_inner.obj", "true true 0 0 0 -1 0.5"),
    "brickwall": ("brickwall.obj", "true true 0.05 0 -1 -10 1"),
}


def unpack(name, root):
    """Write the files of asset `name` into root/name; returns that directory."""
    d = os.path.join(root, name)
    os.makedirs(d, exist_ok=True)
    with np.load(os.path.join(FIXTURES, name + ".npz")) as z:
        for f in z.files:
            path = os.path.join(d, f)
            if not os.path.exists(path):
                z[f].tofile(path)
    return d


def model_line(name, root):
    """The .scene model line for asset `name` (or "african_head/eye")."""
    base = name.split("/")[0]
    f, args = PLACEMENT[name]
    return f"model {os.path.join(unpack(base, root), f)} {args}"


def scene(path, names, mode="forward", width=1280, height=800, shadow="on", ssao="off", ssaa="off 2",
          light="point 2 5 5 1 1 1", camera="persp -1 1 1 0 0 -1", extra=()):
    """A .scene file in the reference's format (src/scene.cpp:19-196) with the
    given assets; returns its path.  The default header is the README's
    console-output run (README.md:190-200): forward, 1280x800, SSAA off,
    shadow on, SSAO off, point light (2,5,5) colour 1, camera (-1,1,1) looking
    at (0,0,-1)."""
    root = os.path.dirname(path)
    lines = [f"mode {mode}", f"screen {width} {height}", f"ssaa {ssaa}", f"ssao {ssao}", f"shadow {shadow}",
             f"light {light}", f"camera {camera}", *extra]
    lines += [model_line(n, root) for n in names]
    os.makedirs(root, exist_ok=True)
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    return path
