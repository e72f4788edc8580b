"""Host formats: the product's C++ .scene/.obj/.mtl/.tga parsers against the
oracle's independent Python restatement, array for array (bit-exact), on
generated scenes and on the reference's own sample assets (when the reference
tree is present; it is not on GPU machines)."""
import os

import numpy as np
import pytest

import forkerrenderer_amd as F
from forkerrenderer_amd import scenegen as G
from oracle import sceneload
from oracle.objload import tga_read

REF = "/root/reference/obj"


def _cmp_models(prod, orc):
    assert len(prod) == len(orc)
    for (pm, pM), (om, oM) in zip(prod, orc):
        for k in ("pos", "uv", "nrm", "vi", "ti", "ni", "face_material"):
            a, b = np.asarray(pm[k]), np.asarray(om[k])
            assert a.shape == b.shape, k
            assert np.array_equal(a.view(np.uint32) if a.dtype == np.float32 else a,
                                  b.view(np.uint32) if b.dtype == np.float32 else b), k
        assert (pm["tan"] is None) == (om["tan"] is None)
        if pm["tan"] is not None:
            assert np.array_equal(pm["tan"].view(np.uint32), om["tan"].view(np.uint32)), "tangents"
        assert pm["support_pbr"] == om["support_pbr"]
        assert np.array_equal(np.asarray(pM, np.float32).view(np.uint32),
                              np.asarray(oM, np.float32).view(np.uint32)), "model matrix"


def _cmp_scene(path):
    sc = F.Scene(path)
    I = sc.info
    prod_models = [sc.model(i) for i in range(I.n_models)]
    o = sceneload.load_scene(path)
    _cmp_models(prod_models, o["models"])
    assert I.n_materials == len(o["materials"])
    for i, om in enumerate(o["materials"]):
        pm = sc.material(i)
        for k, v in om.items():
            pv = getattr(pm, k)
            pv = list(pv) if hasattr(pv, "__len__") else pv
            assert np.allclose(np.float32(pv), np.float32(v), rtol=0, atol=0), (i, k, pv, v)
    assert I.n_textures == len(o["textures"])
    for i, ot in enumerate(o["textures"]):
        pt = sc.texture(i)
        assert np.array_equal(pt["bytes"], ot["bytes"]), i
        assert pt["wrap"] == ot["wrap"] and pt["filter"] == ot["filter"]
    oi = o["info"]
    assert (I.width, I.height, I.ssaa_on, I.ssaa_k, I.ssao_on, I.shadow_on, I.deferred) == \
        (oi["width"], oi["height"], oi["ssaa_on"], oi["ssaa_k"], oi["ssao_on"], oi["shadow_on"], oi["deferred"])
    assert list(I.eye) == oi["eye"] and list(I.look_at) == oi["look_at"]
    assert list(I.lights[0].vec) + list(I.lights[0].color) == oi["point"]
    return sc


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_generated_tiny_scene_loader(scene_dir, seed):
    p = G.tiny_scene(os.path.join(scene_dir, f"ld{seed}"), seed=seed, subdiv=2, tex=8, pbr=seed != 1)
    _cmp_scene(p)


def test_config1_loader(scene_dir):
    p = G.config1(os.path.join(scene_dir, "c1"))
    sc = _cmp_scene(p)
    m, _ = sc.model(1)
    assert len(m["vi"]) == 64 * 32 * 2 - 128  # UV sphere without pole degenerates


def _ref_scene(d, models, extra=""):
    lines = ["mode deferred", "screen 64 48", "light point 2 5 5 1 1 1", "camera persp -1 1 1 0 0 -1", extra]
    for m, norm, gen in models:
        lines.append(f"model {REF}/{m} {norm} {gen} 0 0 -1 30 1")
    os.makedirs(d, exist_ok=True)
    p = os.path.join(d, "ref.scene")
    with open(p, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    return p


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference assets not present on this machine")
@pytest.mark.parametrize("models", [
    [("plane/plane.obj", "false", "false")],
    [("brickwall/brickwall.obj", "true", "true")],
    [("great_sword/great_sword.obj", "true", "true")],   # quads: fan triangulation
    [("mary/mary.obj", "true", "true")],                 # two groups, map order != file order
    [("catbox/catbox.obj", "true", "false")],
    [("african_head/head.obj", "true", "true"), ("african_head/eye_inner.obj", "true", "true")],
], ids=lambda m: m[0][0].split("/")[0])
def test_reference_assets_loader(scene_dir, models):
    p = _ref_scene(os.path.join(scene_dir, "ref_" + models[0][0].split("/")[0]), models, "texture repeat linear")
    _cmp_scene(p)


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference assets not present on this machine")
def test_mary_mesh_order_is_map_order():
    """The reference draws meshes in std::map (name) order, not file order."""
    import re
    txt = open(f"{REF}/mary/mary.obj").read()
    names = re.findall(r"^g (\S+)", txt, flags=re.M)
    assert names != sorted(names), "asset no longer exercises map order"


def test_tga_roundtrip_and_rle_packets(tmp_path):
    rng = np.random.default_rng(0)
    img = rng.integers(0, 4, (37, 53, 3), dtype=np.uint8)  # few values -> runs and raw packets
    img[5:9] = 7
    for rle in (False, True):
        p = str(tmp_path / f"t{int(rle)}.tga")
        F.tga_write(p, img, rle=rle)
        back = tga_read(p)   # oracle decoder, bottom-left origin -> flipped to top-first rows
        assert np.array_equal(back[::-1], img)
        raw = open(p, "rb").read()
        assert raw[2] == (10 if rle else 2) and raw[17] == 0
        assert raw[-18:] == b"TRUEVISION-XFILE.\0"
    grey = rng.integers(0, 255, (9, 11), dtype=np.uint8)
    p = str(tmp_path / "g.tga")
    F.tga_write(p, grey, rle=True)
    assert np.array_equal(tga_read(p)[::-1, :, 0], grey)


def test_rle_packet_structure(tmp_path):
    """RLE stream of a row with a run of 5 equal pixels then 3 distinct ones:
    one run packet (0x80 | 4) and one raw packet of 3 (count-1 = 2)."""
    row = np.array([[[1, 1, 1]] * 5 + [[2, 2, 2], [3, 3, 3], [4, 4, 4]]], np.uint8)
    p = str(tmp_path / "r.tga")
    F.tga_write(p, row, rle=True)
    body = open(p, "rb").read()[18:-26]
    assert body == bytes([0x80 | 4, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 4, 4, 4])


def test_stochastic_scene_keys(scene_dir):
    """`ssaokernel reference` selects Render::DoSSAO with the random stream;
    `shadowfilter refpcf|refpcss` the reference's PCF / PCSS with its
    parameters (shadow.h:18-22: 64 PCF and 32 blocker samples, filter sizes
    0.007 and 0.005, area light 2.5).  A scene without filter keys gets the
    reference's defaults: PCSS and DoSSAO on the random stream, the
    recursive blur."""
    for filt, mode in (("refpcf", F.SHADOW_REF_PCF), ("refpcss", F.SHADOW_REF_PCSS)):
        p = G.tiny_scene(os.path.join(scene_dir, "k" + filt), seed=1, ssao=True,
                         extra=f"ssaokernel reference\nshadowfilter {filt}")
        I = F.Scene(p).info
        assert I.ssao.stochastic == 1
        s = I.shadow
        assert (s.mode, s.n_blocker, s.n_pcf) == (mode, 32, 64)
        assert s.pcf_filter == np.float32(0.007) and s.blocker_filter == 0.005 and s.pcss_filter == 0.007
        assert s.area_light == np.float32(2.5)
    p = G.tiny_scene(os.path.join(scene_dir, "kdet"), seed=1, ssao=True)
    I = F.Scene(p).info
    assert I.ssao.stochastic == 0 and I.shadow.mode == F.SHADOW_HARD
    # no filter keys: the reference's own filters, as its default build has them
    d = os.path.join(scene_dir, "kdefault")
    os.makedirs(d, exist_ok=True)
    p = os.path.join(d, "plain.scene")
    with open(p, "w") as fh:
        fh.write("mode deferred\nssao on\nlight point 2 5 5 1 1 1\ncamera persp -1 1 1 0 0 -1\n")
    I = F.Scene(p).info
    assert (I.shadow_on, I.shadow.mode, I.shadow.n_blocker, I.shadow.n_pcf) == (1, F.SHADOW_REF_PCSS, 32, 64)
    assert I.shadow.blocker_filter == 0.005 and I.shadow.pcss_filter == 0.007 and I.shadow.area_light == np.float32(2.5)
    assert (I.ssao_on, I.ssao.stochastic, I.ssao.blur) == (1, 1, 1)


def test_rle_stream_equals_the_reference_encoder(tmp_path):
    """The product's packetiser (written from the packet rule) against the
    oracle's step-for-step restatement of the reference's encoder loop, on
    rows built to hit every boundary: runs of 1..300 equal pixels (runs
    longer than 128 split), literal stretches of 1..300 (split at 128), a
    literal stretch ending in an equal pair at offsets 125-128, single
    pixels at the image end, 1/3/4 bytes per pixel."""
    from oracle.objload import tga_rle_encode
    rng = np.random.default_rng(5)
    cases = []
    for bpp in (1, 3, 4):
        for _ in range(40):
            parts = []
            for _ in range(int(rng.integers(1, 6))):
                n = int(rng.choice([1, 2, 3, 125, 126, 127, 128, 129, 130, 255, 256, 257, 300]))
                if rng.random() < 0.5:
                    parts.append(np.repeat(rng.integers(0, 256, (1, bpp)), n, 0))
                else:
                    lit = rng.integers(0, 256, (n, bpp))
                    lit[1:, 0] = (lit[:-1, 0] + 1 + lit[1:, 0] % 200) % 256   # neighbours differ
                    parts.append(lit)
            cases.append(np.concatenate(parts).astype(np.uint8))
        cases.append(rng.integers(0, 2, (1000, bpp)).astype(np.uint8))   # dense mix of short runs
        for k in (124, 125, 126, 127, 128):                            # equal pair at offset k
            lit = np.stack([np.arange(300) % 251] * bpp, 1).astype(np.uint8)
            lit[k + 1] = lit[k]
            cases.append(lit)
    for i, px in enumerate(cases):
        bpp = px.shape[1]
        img = px.reshape(1, -1, bpp) if bpp > 1 else px.reshape(1, -1)
        p = str(tmp_path / f"c{i}.tga")
        F.tga_write(p, img, rle=True)
        got = open(p, "rb").read()[18:-26]
        assert got == tga_rle_encode(px.tobytes(), bpp), f"case {i} (bpp {bpp}, {len(px)} px)"


def _asan_tga_driver(tmp_path):
    """host_io.cpp's TGA reader built for the host with AddressSanitizer."""
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = os.path.join(root, "tests", "tga_read_driver.cpp")
    exe = str(tmp_path / "tga_read_asan")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address", "-fno-omit-frame-pointer",
           "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(root, "include"), "-I", "/opt/rocm/include",
           "-I", os.path.join(root, "forkerrenderer_amd", "csrc"), src,
           os.path.join(root, "forkerrenderer_amd", "csrc", "host_io.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip("host build with AddressSanitizer failed: " + r.stderr[-400:])
    return exe


def test_rle_decoder_rejects_packets_past_the_image(tmp_path):
    """A literal packet and a repeat packet that run past the last pixel are
    load errors, raised before any byte lands outside the image (checked
    under AddressSanitizer); a well-formed stream still decodes."""
    import struct
    import subprocess
    exe = _asan_tga_driver(tmp_path)
    env = dict(os.environ, ASAN_OPTIONS="verify_asan_link_order=0:detect_leaks=0")

    def tga(w, h, bpp, body):
        hdr = struct.pack("<BBBHHBHHHHBB", 0, 0, 10 if bpp > 1 else 11, 0, 0, 0, 0, 0, w, h, bpp * 8, 0)
        return hdr + bytes(body)

    cases = {
        "literal_over": (tga(2, 2, 3, [5] + [7] * 18), 1),      # 6 literal pixels into 4
        "repeat_over": (tga(2, 2, 4, [0x80 | 9, 1, 2, 3, 4]), 1),  # 10 repeats into 4
        "repeat_over_grey": (tga(3, 1, 1, [0x80 | 127, 9]), 1),    # 128 repeats into 3
        "ok": (tga(2, 2, 3, [0x80 | 1, 1, 2, 3, 1, 4, 5, 6, 7, 8, 9]), 0),
    }
    for name, (blob, want_err) in cases.items():
        p = tmp_path / f"{name}.tga"
        p.write_bytes(blob)
        r = subprocess.run([exe, str(p)], capture_output=True, text=True, env=env)
        assert "AddressSanitizer" not in r.stderr, name + ": " + r.stderr[-600:]
        if want_err:
            assert r.returncode == 1 and "too many pixels read" in r.stdout, (name, r.stdout, r.stderr[-300:])
        else:
            assert r.returncode == 0, (name, r.stdout, r.stderr[-300:])


def test_texture_modes_are_captured_at_load(scene_dir):
    """Without a `texture` key, textures take the modes ForkerGL holds when
    the scene is loaded (src/forkergl.cpp:15-16, src/model.cpp:425-426):
    NoWrap + Nearest by default, or the modes given to the loader; a
    `texture` key overrides them."""
    p = G.tiny_scene(os.path.join(scene_dir, "texdef"), seed=1, texture=None)
    for args, want in (((), (F.WRAP_NONE, F.FILTER_NEAREST)),
                       ((F.WRAP_MIRROR, F.FILTER_LINEAR), (F.WRAP_MIRROR, F.FILTER_LINEAR))):
        sc = F.Scene(p, *args)
        assert sc.info.n_textures > 0
        assert all((sc.texture(i)["wrap"], sc.texture(i)["filter"]) == want for i in range(sc.info.n_textures))
    p = G.tiny_scene(os.path.join(scene_dir, "texkey"), seed=1, texture="clamp linear")
    sc = F.Scene(p, F.WRAP_REPEAT, F.FILTER_NEAREST)
    assert all((sc.texture(i)["wrap"], sc.texture(i)["filter"]) == (F.WRAP_CLAMP, F.FILTER_LINEAR)
               for i in range(sc.info.n_textures))


# The reference's sample assets from the committed fixtures (tests/golden/assets):
# these run everywhere, the reference tree is not needed.
ASSET_SETS = [["plane", "diablo_pose"], ["plane", "mary"], ["plane", "great_sword"],
              ["african_head", "african_head/eye"], ["brickwall"]]


@pytest.mark.parametrize("names", ASSET_SETS, ids=lambda n: n[-1].split("/")[0])
def test_asset_fixture_loader(scene_dir, names):
    from tests import assets as A
    p = A.scene(os.path.join(scene_dir, "fixtures", names[-1].split("/")[0] + ".scene"), names,
                mode="deferred", width=64, height=48)
    sc = _cmp_scene(p)
    assert sc.info.n_models == len(names)


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference assets not present on this machine")
def test_asset_fixtures_are_the_reference_files():
    from tests import assets as A
    for f in sorted(os.listdir(A.FIXTURES)):
        name = f[:-4]
        with np.load(os.path.join(A.FIXTURES, f)) as z:
            assert z.files
            for k in z.files:
                with open(os.path.join(REF, name, k), "rb") as fh:
                    assert z[k].tobytes() == fh.read(), (name, k)
