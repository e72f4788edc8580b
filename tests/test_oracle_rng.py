"""The oracle's restatement of the reference's random stream
(oracle/ref_rng.cpp; src/utility.h:90-103, src/geometry.h:947-990) pinned
against the C++ standard's known answer, numpy's independent MT19937, the
closed form of libstdc++'s generate_canonical<float>, and the argument
evaluation order the oracle's compiler gives the reference's
Vector3f(Random(..), Random(..), Random(..)) expressions.

The same numpy restatements (`stream_sphere`, `stream_disk`) are what the
device generator is checked against in tests/test_gpu_stochastic.py."""
import numpy as np

import oracle as O

F32 = np.float32


def mt_raw(n, seed=5489, skip=0):
    """numpy's MT19937 seeded like std::mt19937(seed) (init_genrand)."""
    bg = np.random.MT19937(0)
    bg._legacy_seeding(seed)
    if skip:
        bg.random_raw(skip)
    return bg.random_raw(n).astype(np.uint32)


def canonical(u):
    """generate_canonical<float, 24>(mt19937): one 32-bit draw, float(u) / 2^32
    in float, values that round to 1 replaced by nextafter(1, 0)."""
    r = u.astype(F32) * F32(2.0 ** -32)
    return np.where(r >= F32(1), np.nextafter(F32(1), F32(0)), r).astype(F32)


def rand_pm1(u):
    """Random(-1, 1) = min + (max - min) * Random01(), all in float."""
    return (F32(-1) + F32(2) * canonical(u)).astype(F32)


def stream_sphere(raw, n):
    """The first n RandomVectorInUnitSphere results of a raw stream: tries of
    three draws, arguments evaluated right to left (z from the first draw),
    kept when (x*x + y*y) + z*z < 1.  Returns (vectors, draws consumed)."""
    t = len(raw) // 3
    f = rand_pm1(raw[:3 * t]).reshape(t, 3)
    z, y, x = f[:, 0], f[:, 1], f[:, 2]
    ok = ((x * x + y * y) + z * z) < F32(1)
    idx = np.nonzero(ok)[0][:n]
    assert len(idx) == n, "raw stream too short"
    return np.stack([x[idx], y[idx], z[idx]], 1), 3 * (int(idx[-1]) + 1)


def stream_disk(raw, n):
    """The first n RandomVectorInUnitDisk results: Vector3f(Random(-1, 1),
    Random(-1, 1), 0) evaluated right to left (y from the first draw)."""
    t = len(raw) // 2
    f = rand_pm1(raw[:2 * t]).reshape(t, 2)
    y, x = f[:, 0], f[:, 1]
    ok = ((x * x + y * y) + F32(0) * F32(0)) < F32(1)
    idx = np.nonzero(ok)[0][:n]
    assert len(idx) == n, "raw stream too short"
    return np.stack([x[idx], y[idx]], 1), 2 * (int(idx[-1]) + 1)


def test_mt19937_known_answer_and_numpy_stream():
    """[rand.predef]: the 10000th output of a default-constructed mt19937 is
    4123659995; the whole prefix equals numpy's independent MT19937."""
    O.rng_seed(5489)
    raw = O.rng_raw(10000)
    assert int(raw[-1]) == 4123659995
    assert np.array_equal(raw, mt_raw(10000))
    assert O.rng_draws() == 10000
    O.rng_seed(12345)
    assert np.array_equal(O.rng_raw(3000), mt_raw(3000, seed=12345))


def test_uniform_is_generate_canonical_with_the_clamp():
    O.rng_seed(5489)
    u = O.rng_uniform(50000)
    assert np.array_equal(u.view(np.uint32), canonical(mt_raw(50000)).view(np.uint32))
    # Draw 60,571,531 of the default stream is >= 2^32 - 128: float(u)
    # rounds to 2^32 and the result is clamped to the largest float below 1.
    k = 60_571_531
    assert mt_raw(1, skip=k)[0] >= 0xFFFFFF80
    O.rng_seed(5489)
    left = k
    while left:
        m = min(left, 8_000_000)
        O.rng_raw(m)
        left -= m
    assert O.rng_uniform(1)[0] == np.nextafter(F32(1), F32(0))


def test_vector_arguments_are_drawn_right_to_left():
    """The oracle is built by the g++ that builds the reference; for a
    three-argument constructor call it evaluates the arguments right to left,
    so RandomVector3f's z comes from the first draw (and the disk's y)."""
    for seed in range(1, 200):
        raw = mt_raw(3, seed=seed)
        f = rand_pm1(raw)
        if (f[2] * f[2] + f[1] * f[1]) + f[0] * f[0] < F32(1):
            break
    O.rng_seed(seed)
    v = O.rng_sphere(1)[0]
    assert np.array_equal(v, np.array([f[2], f[1], f[0]], F32))
    assert not np.array_equal(v, f)
    assert O.rng_draws() == 3
    for seed in range(1, 200):
        f = rand_pm1(mt_raw(2, seed=seed))
        if f[0] * f[0] + f[1] * f[1] < F32(1) and f[0] != f[1]:
            break
    O.rng_seed(seed)
    assert np.array_equal(O.rng_disk(1)[0], np.array([f[1], f[0]], F32))


def test_unit_sphere_and_disk_rejection_streams():
    """Rejection sampling consumes whole tries; the accepted vectors and the
    draw counts equal the vectorised restatement, and interleaving sphere and
    disk requests continues one stream."""
    raw = mt_raw(400000)
    O.rng_seed(5489)
    sph = O.rng_sphere(20000)
    want, used = stream_sphere(raw, 20000)
    assert np.array_equal(sph, want)
    assert O.rng_draws() == used
    dsk = O.rng_disk(30000)
    want_d, used_d = stream_disk(raw[used:], 30000)
    assert np.array_equal(dsk, want_d)
    assert O.rng_draws() == used + used_d
    acc = len(sph) * 3 / used    # accepted tries / tries
    assert abs(acc - np.pi / 6) < 0.01   # the unit ball's share of the cube


def test_hemisphere_flips_against_the_normal():
    O.rng_seed(99)
    raw = mt_raw(3000, seed=99)
    want, _ = stream_sphere(raw, 200)
    n = np.array([0.3, -0.5, 0.8], F32)
    got = np.empty((200, 3), F32)
    out = np.empty(3, F32)
    for i in range(200):
        O.lib().orc_rng_hemisphere(O._p(n), O._p(out))
        got[i] = out
    d = (F32(0) + want[:, 0] * n[0] + want[:, 1] * n[1] + want[:, 2] * n[2]).astype(F32)
    flip = np.where(d > F32(0), F32(1), F32(-1))[:, None]
    assert np.array_equal(got, (want * flip).astype(F32))
