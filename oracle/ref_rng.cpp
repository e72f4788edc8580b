/*
 * ref_rng.cpp -- the reference's random-number stream, for the oracle.
 *
 * TEST INFRASTRUCTURE ONLY (part of liboracle.so; see fgl_oracle.h).
 *
 * src/utility.h:90-103 draws every random float from one function-static
 * std::mt19937 (default seed 5489) through a function-static
 * std::uniform_real_distribution<float>(0, 1); src/geometry.h:947-990 builds
 * the random vectors from it.  This file uses the same two standard-library
 * types (libstdc++ of the g++ that builds it) and the same expression shapes:
 * Vector3f(Random(..), Random(..), Random(..)) is a three-argument
 * constructor call whose argument evaluation order is unspecified by C++;
 * the g++ that builds the reference evaluates such calls right to left, and
 * so does the same g++ building this file (tests pin the order it produced).
 *
 * The engine is wrapped to count the 32-bit draws consumed (the device path
 * keeps the same count).
 */
#include <cstdint>
#include <cstring>
#include <random>

namespace {

struct CountingEngine {
    typedef std::mt19937::result_type result_type;
    std::mt19937 e;
    uint64_t draws = 0;
    static constexpr result_type min() { return std::mt19937::min(); }
    static constexpr result_type max() { return std::mt19937::max(); }
    result_type operator()()
    {
        ++draws;
        return e();
    }
};

CountingEngine g_engine;
std::uniform_real_distribution<float> g_dist(0.f, 1.f);

struct Vec3 {
    float x, y, z;
    Vec3(float a, float b, float c) : x(a), y(b), z(c) {}
    float LengthSquared() const { return x * x + y * y + z * z; }
};

// Random01 / Random (src/utility.h:90-103)
inline float Random01() { return g_dist(g_engine); }
inline float Random(float lo, float hi) { return lo + (hi - lo) * Random01(); }

// RandomVector3f(min, max) (src/geometry.h:952-955)
inline Vec3 RandomVector3f(float lo, float hi) { return Vec3(Random(lo, hi), Random(lo, hi), Random(lo, hi)); }

// RandomVectorInUnitSphere (src/geometry.h:957-966)
inline Vec3 RandomVectorInUnitSphere()
{
    while (true) {
        Vec3 p = RandomVector3f(-1.f, 1.f);
        if (p.LengthSquared() >= 1.f) continue;
        return p;
    }
}

// RandomVectorInUnitDisk (src/geometry.h:968-976): the integer bounds
// convert to Float, the third constructor argument is a literal 0.
inline Vec3 RandomVectorInUnitDisk()
{
    while (true) {
        Vec3 p = Vec3(Random(-1, 1), Random(-1, 1), 0);
        if (p.LengthSquared() >= 1.f) continue;
        return p;
    }
}

// Dot (src/geometry.h:881-888): from 0.f in index order
inline float Dot(const Vec3 &a, const Vec3 &b)
{
    float r = 0.f;
    r += a.x * b.x;
    r += a.y * b.y;
    r += a.z * b.z;
    return r;
}

}  // namespace

extern "C" {

// Restarts the stream: std::mt19937::seed(seed) (the reference's generator is
// default-constructed, i.e. seed 5489).
void orc_rng_seed(uint32_t seed)
{
    g_engine.e.seed(seed);
    g_engine.draws = 0;
    g_dist.reset();
}

uint64_t orc_rng_draws(void) { return g_engine.draws; }

// n raw 32-bit engine outputs.
void orc_rng_raw(int n, uint32_t *out)
{
    for (int i = 0; i < n; ++i) out[i] = g_engine();
}

// n Random01() values.
void orc_rng_uniform(int n, float *out)
{
    for (int i = 0; i < n; ++i) out[i] = Random01();
}

// n RandomVectorInUnitSphere() results (x, y, z each).
void orc_rng_sphere(int n, float *out)
{
    for (int i = 0; i < n; ++i) {
        Vec3 p = RandomVectorInUnitSphere();
        out[3 * i] = p.x;
        out[3 * i + 1] = p.y;
        out[3 * i + 2] = p.z;
    }
}

// n RandomVectorInUnitDisk() results (x, y each).
void orc_rng_disk(int n, float *out)
{
    for (int i = 0; i < n; ++i) {
        Vec3 p = RandomVectorInUnitDisk();
        out[2 * i] = p.x;
        out[2 * i + 1] = p.y;
    }
}

// RandomVectorInHemisphere(normal) (src/geometry.h:978-990)
void orc_rng_hemisphere(const float n[3], float out[3])
{
    Vec3 v = RandomVectorInUnitSphere();
    if (!(Dot(v, Vec3(n[0], n[1], n[2])) > 0.f)) v = Vec3(-v.x, -v.y, -v.z);
    out[0] = v.x;
    out[1] = v.y;
    out[2] = v.z;
}

}  // extern "C"
