// fgl_math.hpp -- float/double arithmetic of the reference's math library,
// usable from host code and from gfx950 kernels.
//
// Compiled with -ffp-contract=off and without fast-math: each expression is a
// sequence of separately rounded IEEE operations in the reference's order.
//   Dot:      r = 0.f; r += a[i]*b[i] in index order     (src/geometry.h:881-888)
//   v / s:    v * ((float)1 / s)                           (src/geometry.h:336-341)
//   Length:   sqrtf((x*x + y*y) + z*z)                     (src/geometry.h:370-371)
//   Normalize: v * 1 / Length(v)                           (src/geometry.h:906-910)
//   Clamp/Max: std::min/std::max compare semantics         (src/utility.h:31-49)
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define FGL_HD __host__ __device__ __forceinline__

namespace fgl {

struct Vec2 { float x, y; };
struct Vec3 { float x, y, z; };
struct Vec4 { float x, y, z, w; };

FGL_HD Vec2 vec2(float x, float y) { Vec2 r; r.x = x; r.y = y; return r; }
FGL_HD Vec3 vec3(float x, float y, float z) { Vec3 r; r.x = x; r.y = y; r.z = z; return r; }
FGL_HD Vec4 vec4(float x, float y, float z, float w) { Vec4 r; r.x = x; r.y = y; r.z = z; r.w = w; return r; }

// std::max(a,b) == (a < b) ? b : a ; std::min(a,b) == (b < a) ? b : a
FGL_HD float smax(float a, float b) { return (a < b) ? b : a; }
FGL_HD float smin(float a, float b) { return (b < a) ? b : a; }
FGL_HD int imax(int a, int b) { return (a < b) ? b : a; }
FGL_HD int imin(int a, int b) { return (b < a) ? b : a; }
FGL_HD float clampf(float v, float lo, float hi) { return smin(hi, smax(v, lo)); }
FGL_HD int clampi(int v, int lo, int hi) { return imin(hi, imax(v, lo)); }

// 1.f / s, correctly rounded.  On the device, for |s| in [2^-125, 2^125],
// the hardware reciprocal refined by one fma Newton step: equal to the IEEE
// quotient for every such float (all 4,194,304,002 of them are compared by
// fgl_check_recip), in 3 operations instead of the 11 of the full division;
// other s take the division.
FGL_HD float recip(float s)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const float a = fabsf(s);
    if (__builtin_expect(a >= 0x1p-125f && a <= 0x1p125f, 1)) {
        const float y = __builtin_amdgcn_rcpf(s);
        return fmaf(fmaf(-s, y, 1.f), y, y);
    }
#endif
    return 1.f / s;
}

// recip(s) for a caller that has shown 2^-125 <= |s| <= 2^125: on the
// device the reciprocal step alone, without the range test.
FGL_HD float recip_in_range(float s)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const float y = __builtin_amdgcn_rcpf(s);
    return fmaf(fmaf(-s, y, 1.f), y, y);
#else
    return 1.f / s;
#endif
}

// float -> int as x86 cvttss2si: truncation; NaN or outside [-2^31, 2^31)
// gives INT_MIN.  gfx950's v_cvt_i32_f32 saturates instead, so range-check.
FGL_HD int trunc_int(float v)
{
    return (v >= -2147483648.0f && v < 2147483648.0f) ? (int)v : (int)0x80000000;
}

FGL_HD float dot3(Vec3 a, Vec3 b)
{
    float r = 0.f;
    r += a.x * b.x;
    r += a.y * b.y;
    r += a.z * b.z;
    return r;
}
FGL_HD float dot4(Vec4 a, Vec4 b)
{
    float r = 0.f;
    r += a.x * b.x;
    r += a.y * b.y;
    r += a.z * b.z;
    r += a.w * b.w;
    return r;
}
FGL_HD Vec3 add3(Vec3 a, Vec3 b) { return vec3(a.x + b.x, a.y + b.y, a.z + b.z); }
FGL_HD Vec3 sub3(Vec3 a, Vec3 b) { return vec3(a.x - b.x, a.y - b.y, a.z - b.z); }
FGL_HD Vec3 mul3(Vec3 a, Vec3 b) { return vec3(a.x * b.x, a.y * b.y, a.z * b.z); }
FGL_HD Vec3 scale3(Vec3 a, float s) { return vec3(a.x * s, a.y * s, a.z * s); }
FGL_HD Vec3 divs3(Vec3 a, float s) { float inv = recip(s); return scale3(a, inv); }
FGL_HD float len3(Vec3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
FGL_HD Vec3 norm3(Vec3 a) { return divs3(scale3(a, 1.f), len3(a)); }
FGL_HD Vec3 cross3(Vec3 a, Vec3 b)
{
    return vec3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
FGL_HD Vec3 lerp3(float t, Vec3 a, Vec3 b) { return add3(scale3(a, 1 - t), scale3(b, t)); }
FGL_HD Vec3 clamp01_3(Vec3 a)
{
    return vec3(clampf(a.x, 0.f, 1.f), clampf(a.y, 0.f, 1.f), clampf(a.z, 0.f, 1.f));
}
FGL_HD Vec4 scale4(Vec4 a, float s) { return vec4(a.x * s, a.y * s, a.z * s, a.w * s); }
FGL_HD Vec4 divs4(Vec4 a, float s) { float inv = recip(s); return scale4(a, inv); }

// Row-major 4x4 times column vector, one Dot per row (src/geometry.h:774-782).
FGL_HD Vec4 mv4(const float *m, Vec4 v)
{
    Vec4 o;
    o.x = dot4(vec4(m[0], m[1], m[2], m[3]), v);
    o.y = dot4(vec4(m[4], m[5], m[6], m[7]), v);
    o.z = dot4(vec4(m[8], m[9], m[10], m[11]), v);
    o.w = dot4(vec4(m[12], m[13], m[14], m[15]), v);
    return o;
}
FGL_HD Vec3 mv3(const float *m, Vec3 v)
{
    return vec3(dot3(vec3(m[0], m[1], m[2]), v), dot3(vec3(m[3], m[4], m[5]), v),
                dot3(vec3(m[6], m[7], m[8]), v));
}

// Matrix3x3 whose columns are the triangle's three vertex values, times the
// barycentric vector: one Dot per row (src/shaders/phongshader.h:94-97).
FGL_HD Vec3 interp3(Vec3 c0, Vec3 c1, Vec3 c2, Vec3 b)
{
    return vec3(dot3(vec3(c0.x, c1.x, c2.x), b), dot3(vec3(c0.y, c1.y, c2.y), b),
                dot3(vec3(c0.z, c1.z, c2.z), b));
}
FGL_HD Vec2 interp2(Vec2 c0, Vec2 c1, Vec2 c2, Vec3 b)
{
    return vec2(dot3(vec3(c0.x, c1.x, c2.x), b), dot3(vec3(c0.y, c1.y, c2.y), b));
}

// Barycentric (src/geometry.cpp:20-56): FP32 differences of the integer
// window coordinates, widened to FP64 for the cross product and 1.0/c.z,
// narrowed to float.  Returns false for degenerate or outside.
FGL_HD bool barycentric_ref(int ax, int ay, int bx, int by, int cx, int cy, int px, int py,
                            float *b0, float *b1, float *b2)
{
    float Ax = (float)ax, Ay = (float)ay, Bx = (float)bx, By = (float)by;
    float Cx = (float)cx, Cy = (float)cy, Px = (float)px, Py = (float)py;
    double s0x = (double)(Bx - Ax), s0y = (double)(Cx - Ax), s0z = (double)(Ax - Px);
    double s1x = (double)(By - Ay), s1y = (double)(Cy - Ay), s1z = (double)(Ay - Py);
    double ux = s0y * s1z - s0z * s1y;
    double uy = s0z * s1x - s0x * s1z;
    double uz = s0x * s1y - s0y * s1x;
    if (!(fabs(uz) > 1e-2)) return false;
    double inv = 1.0 / uz;
    ux *= inv;
    uy *= inv;
    float r0 = (float)(1.0 - (ux + uy)), r1 = (float)ux, r2 = (float)uy;
    if (r0 < 0.f || r1 < 0.f || r2 < 0.f) return false;
    *b0 = r0; *b1 = r1; *b2 = r2;
    return true;
}

}  // namespace fgl
