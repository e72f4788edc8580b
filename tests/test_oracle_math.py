"""Pins the CPU oracle independently of its own code.

Each check compares an oracle function with something the mathematics fixes
-- exact rational arithmetic, float64 closed forms, a library routine, a
brute-force evaluation in a different loop order, or an invariant -- chosen so
that a plausible mistake (dropped term, wrong sign or index, transposed
operand, wrong rounding step) fails.  Tolerances are stated per check.
"""
import math
from fractions import Fraction

import numpy as np
import pytest
from scipy import ndimage

import oracle as O

INT_MIN = -2 ** 31


def test_trunc_int_matches_x86_cvttss2si():
    lib = O.lib()
    for v in [0.0, -0.0, 0.999, -0.999, 1.5, -1.5, 123456.7, -98765.4, 2147483520.0, -2147483648.0]:
        assert lib.orc_trunc_int(v) == int(np.float32(v))
    for v in [2147483648.0, 3e9, -2147483904.0, -3e9, float("inf"), float("-inf"), float("nan")]:
        assert lib.orc_trunc_int(v) == INT_MIN


def _exact_bary(A, B, C, P):
    s0 = (B[0] - A[0], C[0] - A[0], A[0] - P[0])
    s1 = (B[1] - A[1], C[1] - A[1], A[1] - P[1])
    cx = s0[1] * s1[2] - s0[2] * s1[1]
    cy = s0[2] * s1[0] - s0[0] * s1[2]
    cz = s0[0] * s1[1] - s0[1] * s1[0]
    if cz == 0:
        return None
    u, v = Fraction(cx, cz), Fraction(cy, cz)
    return (1 - u - v, u, v)


def test_barycentric_against_exact_rationals():
    """Coverage equals the exact closed-triangle test except where the exact
    first weight is 0 (then FP64 rounding decides); weights are within 1e-6."""
    rng = np.random.default_rng(1)
    n_in = 0
    for _ in range(4000):
        A, B, C = (tuple(int(x) for x in rng.integers(-40, 40, 2)) for _ in range(3))
        P = tuple(int(x) for x in rng.integers(-45, 45, 2))
        got = O.barycentric(A, B, C, P)
        ex = _exact_bary(A, B, C, P)
        if ex is None:
            assert got is None
            continue
        if any(e < 0 for e in ex[1:]) or (ex[0] < 0):
            assert got is None, (A, B, C, P, ex)
            continue
        if ex[0] == 0:
            if got is not None:
                assert got[0] >= 0
            continue
        assert got is not None, (A, B, C, P, ex)
        n_in += 1
        for g, e in zip(got, ex):
            assert abs(float(g) - float(e)) <= 1e-6 * max(1.0, abs(float(e)))
    assert n_in > 200


def test_shared_edge_covered_by_both_triangles():
    # quad split along the diagonal A-C; P on the diagonal (weight of B or D is 0)
    A, B, C, D = (0, 0), (10, 0), (10, 10), (0, 10)
    for P in [(5, 5), (3, 3), (7, 7)]:
        assert O.barycentric(A, B, C, P) is not None
        assert O.barycentric(A, C, D, P) is not None


def test_degenerate_and_tiny_triangles_are_dropped():
    assert O.barycentric((0, 0), (5, 5), (10, 10), (5, 5)) is None     # collinear
    assert O.barycentric((3, 3), (3, 3), (3, 3), (3, 3)) is None        # a point


def test_matrix_builders_against_float64_closed_forms():
    eye, at = np.array([-1.0, 1.0, 1.0]), np.array([0.0, 0.0, -1.0])
    V = O.make_lookat(eye, at).astype(np.float64)
    f = (at - eye) / np.linalg.norm(at - eye)
    r = np.cross(f, [0, 1, 0]); r /= np.linalg.norm(r)
    u = np.cross(r, f)
    R = np.eye(4); R[0, :3], R[1, :3], R[2, :3] = r, u, -f
    T = np.eye(4); T[:3, 3] = -eye
    assert np.allclose(V, R @ T, atol=1e-6)
    P = O.make_perspective(45.0, 1.6, 0.01, 20.0).astype(np.float64)
    t = math.tan(math.radians(22.5))
    want = np.array([[1 / (1.6 * t), 0, 0, 0], [0, 1 / t, 0, 0],
                     [0, 0, -(20.01) / 19.99, -2 * 20 * 0.01 / 19.99], [0, 0, -1, 1.0]])
    assert np.allclose(P, want, rtol=1e-6, atol=1e-7)   # note m[3][3] == 1 (the reference keeps it)
    Or = O.make_ortho(-3, 3, -2, 2, 0.1, 20).astype(np.float64)
    wantO = np.array([[1 / 3, 0, 0, 0], [0, 0.5, 0, 0], [0, 0, -2 / 19.9, -20.1 / 19.9], [0, 0, 0, 1]])
    assert np.allclose(Or, wantO, rtol=1e-6, atol=1e-7)
    M = O.make_model_matrix((0.5, -1.0, 2.0), 30.0, 1.7).astype(np.float64)
    c, s = math.cos(math.radians(30)), math.sin(math.radians(30))
    Rm = np.array([[c, 0, s, 0], [0, 1, 0, 0], [-s, 0, c, 0], [0, 0, 0, 1]])
    Tm = np.eye(4); Tm[:3, 3] = (0.5, -1.0, 2.0)
    assert np.allclose(M, Tm @ Rm @ np.diag([1.7, 1.7, 1.7, 1.0]), atol=1e-6)
    N = O.make_normal_matrix(M).astype(np.float64)
    assert np.allclose(N, np.linalg.inv(M).T[:3, :3], rtol=1e-5, atol=1e-6)
    VP = O.make_viewport(0, 0, 640, 400).astype(np.float64)
    assert np.allclose(VP, [[320, 0, 0, 320], [0, 200, 0, 200], [0, 0, 0.5, 0.5], [0, 0, 0, 1]])


def test_mat_mul_is_row_column_dot():
    rng = np.random.default_rng(2)
    a, b = rng.normal(size=(4, 4)).astype(np.float32), rng.normal(size=(4, 4)).astype(np.float32)
    assert np.allclose(O.mat_mul(a, b), a.astype(np.float64) @ b.astype(np.float64), atol=1e-5)
    assert not np.allclose(O.mat_mul(a, b), (b.astype(np.float64) @ a.astype(np.float64)), atol=1e-3)


def _tex(img_rgb, wrap, filt):
    bgr = np.ascontiguousarray(img_rgb[..., ::-1])
    t = O.Texture(bgr.shape[1], bgr.shape[0], 3, bgr.ctypes.data, wrap, filt)
    return t, bgr


def _sample(t, u, v):
    out = np.zeros(3, np.float32)
    O.lib().orc_texture_sample(t, u, v, out.ctypes.data_as(O.C.c_void_p))
    return out


def test_texture_nearest_and_bilinear_against_textbook():
    """Nearest: texel floor(u*(W-0.001)); bilinear: texel-centre weights on the
    clamped 2x2 neighbourhood; repeat = fractional part; no-wrap reads black
    outside.  float64 textbook evaluation, tolerance 2e-6 (in 0..1 units)."""
    rng = np.random.default_rng(3)
    img = rng.integers(0, 256, (5, 4, 3)).astype(np.uint8)
    H, W = img.shape[:2]
    for wrap in (0, 1, 3):
        for filt in (0, 1):
            t, keep = _tex(img, wrap, filt)
            for _ in range(300):
                u, v = rng.uniform(-1.5, 2.5, 2)
                if wrap == 1:
                    uu, vv = u - math.floor(u), v - math.floor(v)
                elif wrap == 3:
                    uu, vv = min(max(u, 0), 1), min(max(v, 0), 1)
                else:
                    uu, vv = u, v
                w_, h_ = W - 0.001, H - 0.001

                def texel(x, y):
                    if wrap != 0:
                        x, y = min(max(x, 0), W - 1), min(max(y, 0), H - 1)
                    if 0 <= x < W and 0 <= y < H:
                        return img[y, x].astype(np.float64)
                    return np.zeros(3)
                if filt == 0:
                    want = texel(math.floor(uu * w_), math.floor(vv * h_))
                else:
                    px, py = uu * w_, vv * h_
                    x0, y0 = math.floor(px - 0.5), math.floor(py - 0.5)
                    tx, ty = px - (x0 + 0.5), py - (y0 + 0.5)
                    top = texel(x0, y0) * (1 - tx) + texel(x0 + 1, y0) * tx
                    bot = texel(x0, y0 + 1) * (1 - tx) + texel(x0 + 1, y0 + 1) * tx
                    want = top * (1 - ty) + bot * ty
                got = _sample(t, np.float32(u), np.float32(v))
                assert np.allclose(got, want / 255.0, atol=2e-6), (wrap, filt, u, v, got, want / 255)


def _fp64_blinn_phong(L, Hh, N, vis, diff, em, param, lc, shadow_on):
    dl, el = np.power(diff, 2.2), np.power(em, 2.2)
    ao, ks, sh = param
    d = max(0.0, float(np.dot(L, N)))
    sp = max(0.0, float(np.dot(Hh, N))) ** sh
    amb, dif, spe = 0.3 * dl * ao, dl * d * ao, ks * sp * np.ones(3)
    if shadow_on:
        f = 1 - (1 - vis) * 0.6
        dif, spe = dif * f, spe * f
    c = amb + (dif + spe + el) * lc
    return np.clip(np.power(c / (c + 1), 1 / 2.2), 0, 1)


def _fp64_pbr(L, V, Hh, N, vis, al, em, param, rad, shadow_on):
    ao, m, r = param
    al, el = np.power(al, 2.2), np.power(em, 2.2)
    nv, nl = max(float(np.dot(N, V)), 0), max(float(np.dot(N, L)), 0)
    nh, hv = max(float(np.dot(N, Hh)), 0), max(float(np.dot(Hh, V)), 0)
    F0 = 0.04 * (1 - m) + al * m
    a2 = (r * r) ** 2
    D = a2 / math.pi / (nh * nh * (a2 - 1) + 1) ** 2
    k = (r + 1) ** 2 / 8
    G = (nv / (nv * (1 - k) + k)) * (nl / (nl * (1 - k) + k))
    Fr = F0 + (1 - F0) * max(1 - hv, 0) ** 5
    spec = Fr * D * G / (4 * nv * nl + 0.001)
    kd = (1 - Fr) * (1 - m)
    Lo = (kd * al / math.pi + spec) * rad * nl
    if shadow_on:
        Lo = Lo * (1 - (1 - vis) * 0.6)
    c = Lo + 0.3 * al * ao + el
    return np.clip(np.power(c / (c + 1), 1 / 2.2), 0, 1)


def _unit(rng):
    v = rng.normal(size=3)
    return v / np.linalg.norm(v)


def test_lighting_against_fp64_reevaluation():
    """FP32 shading error against FP64 re-evaluation of the lighting
    equations: Blinn-Phong and Cook-Torrance (GGX D, Smith/Schlick-GGX G,
    Fresnel-Schlick F).  Stated bound: 1e-5 absolute on [0,1] outputs."""
    rng = np.random.default_rng(4)
    P = O.C.c_void_p
    worst = {"blinn_phong": 0.0, "pbr": 0.0}
    for _ in range(2000):
        N = _unit(rng); L = _unit(rng); V = _unit(rng)
        Hh = (L + V) / np.linalg.norm(L + V)
        vis = float(rng.uniform())
        al = rng.uniform(0, 1, 3); em = rng.uniform(0, 0.2, 3); lc = rng.uniform(0.5, 3, 3)
        a32 = [np.asarray(x, np.float32) for x in (L, Hh, N, al, em, lc)]
        param = np.array([rng.uniform(0.5, 1), rng.uniform(0, 1), rng.uniform(1, 60)], np.float32)
        out = np.zeros(3, np.float32)
        O.lib().orc_shade_blinn_phong(*(x.ctypes.data_as(P) for x in a32[:3]), np.float32(vis),
                                      a32[3].ctypes.data_as(P), a32[4].ctypes.data_as(P),
                                      param.ctypes.data_as(P), a32[5].ctypes.data_as(P), 1,
                                      out.ctypes.data_as(P))
        want = _fp64_blinn_phong(*(x.astype(np.float64) for x in a32[:3]), np.float32(vis),
                                 a32[3].astype(np.float64), a32[4].astype(np.float64),
                                 param.astype(np.float64), a32[5].astype(np.float64), True)
        worst["blinn_phong"] = max(worst["blinn_phong"], float(np.abs(out - want).max()))
        pp = np.array([rng.uniform(0.5, 1), rng.uniform(0, 1), rng.uniform(0.05, 1)], np.float32)
        v32 = np.asarray(V, np.float32)
        O.lib().orc_shade_pbr(a32[0].ctypes.data_as(P), v32.ctypes.data_as(P), a32[1].ctypes.data_as(P),
                              a32[2].ctypes.data_as(P), np.float32(vis), a32[3].ctypes.data_as(P),
                              a32[4].ctypes.data_as(P), pp.ctypes.data_as(P), a32[5].ctypes.data_as(P), 1,
                              out.ctypes.data_as(P))
        want = _fp64_pbr(a32[0].astype(np.float64), v32.astype(np.float64), a32[1].astype(np.float64),
                         a32[2].astype(np.float64), np.float32(vis), a32[3].astype(np.float64),
                         a32[4].astype(np.float64), pp.astype(np.float64), a32[5].astype(np.float64), True)
        worst["pbr"] = max(worst["pbr"], float(np.abs(out - want).max()))
    print("FP32 vs FP64 lighting max abs error:", worst)
    assert worst["blinn_phong"] < 1e-5 and worst["pbr"] < 1e-5, worst


def test_separable_blur_equals_library_convolution():
    rng = np.random.default_rng(5)
    a = rng.uniform(0, 1, (23, 37)).astype(np.float32)
    sigma, R = 2.0, 6
    w = np.exp(-np.arange(R + 1) ** 2 / (2 * sigma * sigma))
    w = (w / (w[0] + 2 * w[1:].sum())).astype(np.float32)
    got = O.blur_separable(a.reshape(-1), 37, 23, w, R).reshape(23, 37)
    k = np.concatenate([w[:0:-1], w]).astype(np.float64)
    want = ndimage.convolve1d(ndimage.convolve1d(a.astype(np.float64), k, axis=1, mode="nearest"),
                              k, axis=0, mode="nearest")
    assert np.allclose(got, want, atol=2e-6)


def test_recursive_blur_is_in_place():
    """Buffer1f::TwoPassGaussianBlurDenoised reads already-blurred left/lower
    neighbours.  Checked against a float64 sequential evaluation and against
    the non-recursive result (which must differ)."""
    rng = np.random.default_rng(6)
    H, W = 9, 14
    a = rng.uniform(0, 1, (H, W))
    w = [0.227027, 0.1945946, 0.1216216, 0.054054, 0.016216]
    b = a.copy()
    for h in range(H):
        for x in range(W):
            r = b[h, x] * w[0]
            for i in range(1, 5):
                r += b[h, min(x + i, W - 1)] * w[i] + b[h, max(x - i, 0)] * w[i]
            b[h, x] = r
    for h in range(H):
        for x in range(W):
            r = b[h, x] * w[0]
            for i in range(1, 5):
                r += b[min(h + i, H - 1), x] * w[i] + b[max(h - i, 0), x] * w[i]
            b[h, x] = r
    got = O.blur_recursive(a.astype(np.float32).reshape(-1), W, H).reshape(H, W)
    assert np.allclose(got, b, atol=2e-6)
    k = np.array(w[:0:-1] + w)
    sep = ndimage.convolve1d(ndimage.convolve1d(a, k, axis=1, mode="nearest"), k, axis=0, mode="nearest")
    assert np.abs(got - sep).max() > 1e-3
    const = O.blur_recursive(np.full(W * H, 0.5, np.float32), W, H)
    assert np.allclose(const, 0.5 * (w[0] + 2 * sum(w[1:])) ** 2, atol=1e-6)


def test_quantise_and_ssaa_resolve_closed_form():
    rng = np.random.default_rng(7)
    W, H, k = 12, 8, 2
    fb = rng.uniform(0, 1, (3, W * H)).astype(np.float32)
    q = O.quantize_rgb(fb, W, H).reshape(H, W, 3)
    want_q = np.floor(fb.astype(np.float64) * np.float64(np.float32(254.99))).astype(np.int64)
    assert np.array_equal(q[..., 2].reshape(-1), want_q[0]) and np.array_equal(q[..., 0].reshape(-1), want_q[2])
    img = O.ssaa_resolve(q.reshape(-1), W, H, k).reshape(H // k, W // k, 3)
    s = q.reshape(H // k, k, W // k, k, 3).astype(np.int64).sum(axis=(1, 3))
    assert np.array_equal(img, s // (k * k))


def _draw(pos, faces, model=np.eye(4, dtype=np.float32)):
    pos = np.asarray(pos, np.float32)
    n = len(pos)
    md = {"pos": pos, "uv": np.zeros((n, 2), np.float32), "nrm": np.tile([[0, 0, 1]], (n, 1)).astype(np.float32),
          "tan": None, "vi": np.asarray(faces, np.int32), "ti": np.asarray(faces, np.int32),
          "ni": np.asarray(faces, np.int32), "face_material": np.zeros(len(faces), np.int32), "support_pbr": False}
    return md, model


def test_depth_fold_brute_force_in_pixel_order():
    """The shadow pass's depth buffer equals a per-pixel fold over triangles in
    submission order (pixel-outer loop, the opposite nesting of the oracle),
    with the reference's reject-if-greater-or-equal test."""
    rng = np.random.default_rng(8)
    W, H = 14, 11
    pos = rng.uniform(-1.2, 1.2, (30, 3)).astype(np.float32)
    pos[:, 2] = rng.uniform(-0.9, 0.9, 30)
    faces = rng.integers(0, 30, (25, 3))
    faces[5] = faces[4]   # an exact duplicate: equal depths, first wins
    md, M = _draw(pos, faces)
    sc = O.OracleScene([(md, M)], [O.Material()], [])
    vp = O.make_viewport(0, 0, W, H)
    fr = O.make_frame(np.eye(4), np.eye(4), np.eye(4), vp, vp, [0, 0, 0], 1, None, [])
    depth, _ = O.shadow_pass(sc, fr, W, H)
    ndc = pos  # L*M = I, w = 1
    win = [(O.lib().orc_trunc_int(np.float32(np.float32(W / 2) * x + np.float32(W / 2))),
            O.lib().orc_trunc_int(np.float32(np.float32(H / 2) * y + np.float32(H / 2))),
            np.float32(np.float32(0.5) * z + np.float32(0.5))) for x, y, z in ndc]
    want = np.full((H, W), np.finfo(np.float32).max, np.float32)
    for py in range(H):
        for px in range(W):
            for f in faces:
                A, B, C = win[f[0]], win[f[1]], win[f[2]]
                xs, ys = [A[0], B[0], C[0]], [A[1], B[1], C[1]]
                if not (min(max(min(xs), 0), W - 1) <= px <= min(max(max(xs), 0), W - 1)):
                    continue
                if not (min(max(min(ys), 0), H - 1) <= py <= min(max(max(ys), 0), H - 1)):
                    continue
                b = O.barycentric(A[:2], B[:2], C[:2], (px, py))
                if b is None:
                    continue
                d = np.float32(np.float32(np.float32(np.float32(0) + b[0] * A[2]) + b[1] * B[2]) + b[2] * C[2])
                if not (d >= want[py, px]):
                    want[py, px] = d
    assert np.array_equal(depth.reshape(H, W).view(np.uint32), want.view(np.uint32))


def test_equal_depth_tie_goes_to_first_submitted():
    quad = [(-1, -1, 0), (1, -1, 0), (1, 1, 0), (-1, 1, 0)]
    faces = [(0, 1, 2), (0, 2, 3)]
    mats = []
    for c in ((1, 0, 0), (0, 1, 0)):
        m = O.Material()
        m.kd[:] = c
        m.map_kd = m.map_ks = m.map_ke = m.map_bump = -1
        m.map_base = m.map_rough = m.map_metal = m.map_ao = m.map_pbr_ke = -1
        mats.append(m)
    d0, M = _draw(quad, faces)
    d1 = dict(d0)
    d1["face_material"] = np.ones(2, np.int32)
    W, H = 16, 12
    vp = O.make_viewport(0, 0, W, H)
    fr = O.make_frame(np.eye(4), np.eye(4), np.eye(4), vp, vp, [0, 0, 0], 0, None, [])
    for order, first in (((d0, d1), 0), ((d1, d0), 1)):
        sc = O.OracleScene([(order[0], M), (order[1], M)], mats, [])
        depth, g = O.geometry_pass(sc, fr, W, H)
        covered = depth < 3e38
        assert covered.sum() > 100
        alb = g[9:12][:, covered]
        assert np.all(alb[first] == 1.0) and np.all(alb[1 - first] == 0.0)


def test_kernel_ssao_flat_facing_plane_is_unoccluded():
    """A plane facing the camera: every hemisphere tap lies in front of it, so
    AO is exactly 1."""
    W, H = 40, 30
    quad = [(-3, -3, -2), (3, -3, -2), (3, 3, -2), (-3, 3, -2)]
    md, M = _draw(quad, [(0, 1, 2), (0, 2, 3)])
    m = O.Material(); m.map_kd = m.map_ks = m.map_ke = m.map_bump = -1
    m.map_base = m.map_rough = m.map_metal = m.map_ao = m.map_pbr_ke = -1
    sc = O.OracleScene([(md, M)], [m], [])
    V = O.make_lookat([0, 0, 0.5], [0, 0, -1])
    P = O.make_perspective(45.0, W / H, 0.01, 20.0)
    vp = O.make_viewport(0, 0, W, H)
    fr = O.make_frame(V, P, np.eye(4), vp, vp, [0, 0, 0.5], 0, None, [])
    depth, g = O.geometry_pass(sc, fr, W, H)

    class Cfg:
        pass
    rng = np.random.default_rng(9)
    cfg = Cfg()
    k = rng.normal(size=(16, 3)); k[:, 2] = np.abs(k[:, 2]); k /= np.linalg.norm(k, axis=1, keepdims=True)
    cfg.kernel = (k * rng.uniform(0.1, 1, (16, 1))).astype(np.float32).reshape(-1).tolist() + [0.0] * (192 - 48)
    cfg.noise = np.concatenate([rng.uniform(-1, 1, (16, 2)), np.zeros((16, 1))], 1).astype(np.float32).reshape(-1).tolist()
    cfg.n_kernel, cfg.radius, cfg.bias, cfg.range = 16, 0.3, 0.0005, 0.01
    ao = O.ssao_kernel(fr, W, H, depth, g, cfg)
    assert np.all(ao[depth < 3e38] == 1.0)


def _ssao_cfg(n, seed):
    class Cfg:
        pass
    rng = np.random.default_rng(seed)
    cfg = Cfg()
    k = rng.normal(size=(n, 3)); k[:, 2] = np.abs(k[:, 2]); k /= np.linalg.norm(k, axis=1, keepdims=True)
    cfg.kernel = (k * rng.uniform(0.1, 1, (n, 1))).astype(np.float32).reshape(-1).tolist() + [0.0] * (192 - 3 * n)
    cfg.noise = np.concatenate([rng.uniform(-1, 1, (16, 2)), np.zeros((16, 1))], 1).astype(np.float32).reshape(-1).tolist()
    cfg.n_kernel, cfg.radius, cfg.bias, cfg.range = n, 0.5, 0.0005, 0.01
    return cfg


def test_kernel_ssao_matches_direct_fp64_projection():
    """The oracle's linearised kernel SSAO (clip = c0 + r (kx cT + ky cB + kz
    cN), fmaf, window = diagonal * clip / w + offset) against the textbook
    definition evaluated in float64: sample S = P + r (T kx + B ky + N kz),
    window = Viewport * (VP * (S, 1)) / w, the same gather and tests.  Only
    samples within rounding of a threshold may differ; a dropped term, wrong
    sign or swapped axis changes most pixels.  The scene is a concave corner
    (floor and wall), so many taps are occluded."""
    W, H = 48, 36
    floor = [(-3, -1, 1), (3, -1, 1), (3, -1, -4), (-3, -1, -4)]
    wall = [(-3, -1, -2), (3, -1, -2), (3, 3, -2), (-3, 3, -2)]
    md0, M = _draw(floor, [(0, 1, 2), (0, 2, 3)])
    md1, _ = _draw(wall, [(0, 1, 2), (0, 2, 3)])
    m = O.Material(); m.map_kd = m.map_ks = m.map_ke = m.map_bump = -1
    m.map_base = m.map_rough = m.map_metal = m.map_ao = m.map_pbr_ke = -1
    sc = O.OracleScene([(md0, M), (md1, M)], [m], [])
    V = O.make_lookat([0.3, 0.8, 1.5], [0, -0.5, -2])
    P = O.make_perspective(45.0, W / H, 0.01, 20.0)
    vp = O.make_viewport(0, 0, W, H)
    fr = O.make_frame(V, P, np.eye(4), vp, vp, [0.3, 0.8, 1.5], 0, None, [])
    depth, g = O.geometry_pass(sc, fr, W, H)
    cfg = _ssao_cfg(32, 4)
    ao = O.ssao_kernel(fr, W, H, depth, g, cfg).reshape(H, W)
    VP = O.mat_mul(P, V).astype(np.float64).reshape(4, 4)
    VPt = np.asarray(vp, np.float64).reshape(4, 4)
    k = np.array(cfg.kernel[:96], np.float64).reshape(32, 3)
    nz = np.array(cfg.noise, np.float64).reshape(16, 3)
    d2 = depth.reshape(H, W)
    cov = d2 < 3e38
    diff = []
    occluded = 0
    for y in range(H):
        for x in range(W):
            if not cov[y, x]:
                continue
            i = x + y * W
            N = g[0:3, i].astype(np.float64)
            Pw = g[3:6, i].astype(np.float64)
            r = nz[(x & 3) + 4 * (y & 3)]
            T = r - N * (r @ N); T /= np.linalg.norm(T)
            B = np.cross(N, T)
            cnt = 0
            for s in range(32):
                S = Pw + cfg.radius * (T * k[s, 0] + B * k[s, 1] + N * k[s, 2])
                c = VP @ np.append(S, 1.0)
                w = VPt @ (c / c[3])
                sx, sy = math.trunc(w[0]), math.trunc(w[1])
                if sx < 0 or sy < 0 or sx >= W or sy >= H:
                    continue
                cached = float(d2[sy, sx])
                if w[2] >= cached + cfg.bias and abs(float(d2[y, x]) - cached) < cfg.range:
                    cnt += 1
            want = (1.0 - cnt / 32) ** 3
            got_cnt = round(32 * (1 - float(ao[y, x]) ** (1 / 3)))
            occluded += cnt > 0
            diff.append(abs(got_cnt - cnt))
    diff = np.array(diff)
    assert occluded > 0.3 * len(diff)
    assert np.mean(diff == 0) > 0.97 and diff.max() <= 2, (np.mean(diff == 0), diff.max())


def _fold_in_order(init, frags):
    """src/forkergl.cpp:180-199 at one pixel: fragments (g, depth) in
    submission order, replace when !(d >= best)."""
    best, win = np.float32(init), -1
    for g, d in frags:
        if not (d >= best):
            best, win = d, g
    return best, win


def _depth_order(d):
    u = int(np.float32(0.0 if d == 0 else d).view(np.uint32))
    return (~u & 0xFFFFFFFF) if u & 0x80000000 else (u | 0x80000000)


def _fold_any_order(init, frags):
    """The rasteriser's order-free statement of the same fold: key minima
    over (order(depth), g + 1), restricted to the fragments after the last
    NaN fragment when the pixel has one (or when the stored depth is NaN)."""
    nan_g = max([g for g, d in frags if np.isnan(d)], default=-1)
    if nan_g >= 0 or np.isnan(init):
        keys = [(_depth_order(d), g + 1, d) for g, d in frags if g > nan_g and not np.isnan(d)]
    else:
        keys = [(_depth_order(init), 0, np.float32(init))] + [(_depth_order(d), g + 1, d) for g, d in frags]
    if keys:
        _, g1, d = min(keys, key=lambda k: (k[0], k[1]))
        if g1 > 0:
            return d, g1 - 1
    if nan_g >= 0:
        return dict(frags)[nan_g], nan_g
    return np.float32(init), -1


def test_order_free_depth_fold_equals_the_in_order_fold():
    """Random fragment lists over a value set with ties, signed zeros,
    infinities and NaNs, random initial depths (NaN included), shuffled ids:
    the order-free rule picks the same winner and the same depth bits as the
    in-order fold, whatever order the fragments arrive in."""
    rng = np.random.default_rng(31)
    vals = np.array([-1.0, -0.0, 0.0, 0.25, 0.5, 1.0, np.inf, -np.inf, np.nan, -np.nan, np.finfo(np.float32).max],
                    np.float32)
    assert vals[9].view(np.uint32) >> 31 == 1   # a NaN with the sign bit set
    nan_wins = 0
    for _ in range(20000):
        n = int(rng.integers(0, 8))
        ids = np.sort(rng.choice(40, n, replace=False))
        frags = [(int(g), vals[rng.integers(0, len(vals))]) for g in ids]
        init = vals[rng.integers(0, len(vals))] if rng.random() < 0.5 else np.finfo(np.float32).max
        want_d, want_g = _fold_in_order(init, frags)
        shuffled = [frags[i] for i in rng.permutation(n)]
        got_d, got_g = _fold_any_order(init, shuffled)
        assert got_g == want_g, (init, frags)
        if np.isnan(want_d):
            assert np.isnan(got_d)
            nan_wins += 1
        else:
            assert np.float32(got_d).view(np.uint32) == np.float32(want_d).view(np.uint32), (init, frags)
    assert nan_wins > 100
