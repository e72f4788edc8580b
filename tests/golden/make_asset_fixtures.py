"""Write tests/golden/assets/<name>.npz: the reference's sample assets
(obj/<name>/ of forkercat/ForkerRenderer: .obj, .mtl and .tga files, byte
for byte) as data fixtures, so the parity tests can load them with the
product's own .scene/.obj/.mtl/.tga parsers where the reference tree is not
present.  Each npz maps a file name to a uint8 array of that file's bytes
(deflate-compressed).  The models' authors are credited in the reference's
README.md:246-256.

Files the reference's loader never opens are left out: mary's .png (its
material names a .tga that does not exist, so Model::loadTexture warns and the
material draws with Kd alone, src/model.cpp:410-428).

Run in a tree where the reference is present:
    python tests/golden/make_asset_fixtures.py [/root/reference]
"""
import os
import sys

import numpy as np

ASSETS = {
    # name: (directory under obj/, files)
    "diablo_pose": ("diablo_pose", ["diablo_pose.obj", "diablo_pose.mtl", "diablo_pose_diff.tga",
                                    "diablo_pose_spec.tga", "diablo_pose_nm.tga"]),
    "great_sword": ("great_sword", ["great_sword.obj", "great_sword.mtl", "great_sword_diff.tga"]),
    "mary": ("mary", ["mary.obj", "mary.mtl"]),
    "african_head": ("african_head", ["head.obj", "head.mtl", "head_diff.tga", "head_spec.tga", "head_nm.tga",
                                      "eye_inner.obj", "eye_inner.mtl", "eye_inner_diff.tga",
                                      "eye_inner_spec.tga", "eye_inner_nm.tga"]),
    "plane": ("plane", ["plane.obj", "plane.mtl"]),
    "brickwall": ("brickwall", ["brickwall.obj", "brickwall.mtl", "brickwall_diff.tga", "brickwall_spec.tga",
                                "brickwall_nm.tga"]),
}

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "assets")


def main(ref_root):
    os.makedirs(OUT, exist_ok=True)
    for name, (sub, files) in ASSETS.items():
        arrays = {}
        for f in files:
            with open(os.path.join(ref_root, "obj", sub, f), "rb") as fh:
                arrays[f] = np.frombuffer(fh.read(), np.uint8)
        path = os.path.join(OUT, name + ".npz")
        np.savez_compressed(path, **arrays)
        print(f"{path}: {len(files)} files, {os.path.getsize(path)} bytes")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
