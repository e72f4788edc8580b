// forkergl_amd.hpp -- C++ host facade over the C ABI in fgl.h with the names
// the reference's frame driver uses (src/forkergl.h, src/shaders/*.h,
// src/scene.h, src/camera.h, src/geometry.h), so pass code written against
// the reference compiles against this header inside `using namespace fglx;`
// and drives the GPU path instead.
//
//   reference                                   facade (namespace fglx)
//   ForkerGL::TextureWrapMode(Texture::Repeat)  the same; applies to scenes loaded afterwards
//   Scene scene(path)                           Scene scene(path)  (loads and uploads)
//   ForkerGL::InitShadowBuffer(w, h)            the same
//   DepthShader s; s.uModelMatrix = M; ...      the same; uNormalMatrix reaches the vertex stage
//   model.Render(s)                             the same (model: scene.GetModel(i))
//   ForkerGL::DrawScreenSpacePixels(scene)      the same
//   ForkerGL::DepthBuffer.GetValue(x, y)        the same (each access reads the target back;
//                                               .Read() gives a whole host copy)
//
// Matrix4x4f / Matrix3x3f are row-major; products and builders call the
// library's host matrix code (fgl_mat_mul, fgl_make_*), the arithmetic the
// GPU path and the oracle use.  Errors become std::runtime_error.  Faces are
// submitted per model in the reference's order (Model::Render); there is no
// per-face DrawTriangle.
#pragma once

#include <array>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "fgl.h"

namespace fglx {

using Float = float;

inline void check(fgl_ctx *c, int rc, const char *what)
{
    if (rc != FGL_OK) throw std::runtime_error(std::string(what) + ": " + (c ? fgl_last_error(c) : "no context"));
}

// ---------------- geometry types (src/geometry.h) ----------------

struct Vector3f {
    Float x = 0.f, y = 0.f, z = 0.f;
    Vector3f() = default;
    explicit Vector3f(Float v) : x(v), y(v), z(v) {}
    Vector3f(Float a, Float b, Float c) : x(a), y(b), z(c) {}
    Float &operator[](int i) { return (&x)[i]; }
    Float operator[](int i) const { return (&x)[i]; }
    const Float *data() const { return &x; }
};
using Point3f = Vector3f;
using Color3 = Vector3f;

struct Matrix3x3f {
    std::array<Float, 9> m{1, 0, 0, 0, 1, 0, 0, 0, 1};
    Float &operator()(int r, int c) { return m[3 * r + c]; }
    Float operator()(int r, int c) const { return m[3 * r + c]; }
    const Float *data() const { return m.data(); }
};

struct Matrix4x4f {
    std::array<Float, 16> m{1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    Matrix4x4f() = default;
    explicit Matrix4x4f(const Float *rows) { std::memcpy(m.data(), rows, sizeof(m)); }
    static Matrix4x4f Identity() { return Matrix4x4f(); }
    Float &operator()(int r, int c) { return m[4 * r + c]; }
    Float operator()(int r, int c) const { return m[4 * r + c]; }
    const Float *data() const { return m.data(); }
    Float *data() { return m.data(); }
};
using Matrix4f = Matrix4x4f;

// Matrix * Matrix: row . column Dot products (src/geometry.h:784-793).
inline Matrix4x4f operator*(const Matrix4x4f &a, const Matrix4x4f &b)
{
    Matrix4x4f r;
    fgl_mat_mul(a.data(), b.data(), r.data());
    return r;
}

inline Matrix4x4f MakeModelMatrix(const Vector3f &translation, Float yRotate = 0.f, Float scale = 1.f)
{
    Matrix4x4f r;
    fgl_make_model_matrix(translation.x, translation.y, translation.z, yRotate, scale, r.data());
    return r;
}
inline Matrix4x4f MakeLookAtMatrix(const Vector3f &eyePos, const Vector3f &center,
                                   const Vector3f &up = Vector3f(0.f, 1.f, 0.f))
{
    Matrix4x4f r;
    fgl_make_lookat(eyePos.data(), center.data(), up.data(), r.data());
    return r;
}
inline Matrix4x4f MakePerspectiveMatrix(Float fov, Float aspectRatio, Float n, Float f)
{
    Matrix4x4f r;
    fgl_make_perspective(fov, aspectRatio, n, f, r.data());
    return r;
}
inline Matrix4x4f MakeOrthographicMatrix(Float l, Float r, Float b, Float t, Float n, Float f)
{
    Matrix4x4f o;
    fgl_make_ortho(l, r, b, t, n, f, o.data());
    return o;
}
inline Matrix3x3f MakeNormalMatrix(const Matrix4x4f &m)
{
    Matrix3x3f r;
    fgl_make_normal_matrix(m.data(), r.m.data());
    return r;
}

// ---------------- textures (src/texture.h:15-27) ----------------

struct Texture {
    enum WrapMode { NoWrap = FGL_WRAP_NONE, Repeat = FGL_WRAP_REPEAT, MirroredRepeat = FGL_WRAP_MIRROR,
                    ClampToEdge = FGL_WRAP_CLAMP };
    enum FilterMode { Nearest = FGL_FILTER_NEAREST, Linear = FGL_FILTER_LINEAR };
};

// ---------------- lights, camera (src/light.h, src/camera.h) ----------------

struct PointLight {
    Point3f position{0.f, 0.f, 0.f};
    Color3 color{1.f, 1.f, 1.f};
};

class Camera {
public:
    enum ProjectionType { Orthographic, Perspective };
    Camera() = default;
    explicit Camera(const Point3f &lookFrom, const Point3f &lookAt = Point3f(0.f, 0.f, 0.f))
        : eye_(lookFrom), at_(lookAt) {}
    Point3f GetPosition() const { return eye_; }
    Point3f GetLookAt() const { return at_; }
    Matrix4x4f GetViewMatrix() const { return MakeLookAtMatrix(eye_, at_, Vector3f(0.f, 1.f, 0.f)); }
    Matrix4x4f GetPerspectiveMatrix(Float fov, Float aspectRatio, Float n, Float f) const
    {
        return MakePerspectiveMatrix(fov, aspectRatio, n, f);
    }
    Matrix4x4f GetOrthographicMatrix(Float l, Float r, Float b, Float t, Float n, Float f) const
    {
        return MakeOrthographicMatrix(l, r, b, t, n, f);
    }

private:
    Point3f eye_, at_;
};

// ---------------- shader uniform blocks (src/shaders/*.h) ----------------
// The vertex and fragment programs run on the GPU; the uniforms keep the
// reference's member names.

struct DepthShader {
    Matrix4x4f uModelMatrix, uLightSpaceMatrix;
};
struct GShader {
    Matrix4x4f uModelMatrix, uViewMatrix, uProjectionMatrix, uLightSpaceMatrix;
    Matrix3x3f uNormalMatrix;
};
struct BlinnPhongShader {
    Matrix4x4f uModelMatrix, uViewMatrix, uProjectionMatrix, uLightSpaceMatrix;
    Matrix3x3f uNormalMatrix;
    PointLight uPointLight;
    Point3f uEyePos;
};
// PBRShader has Blinn-Phong's uniforms (src/shaders/pbrshader.h:20-30); the
// model's materials pick the lighting model (Model::SupportPBR).
struct PBRShader : BlinnPhongShader {};

struct ForkerGL;
class Scene;

// A model uploaded into the facade's context (Model::Load's product).
class Model {
public:
    Model() = default;
    Model(int id, bool pbr) : id_(id), pbr_(pbr) {}
    int Id() const { return id_; }
    bool SupportPBR() const { return pbr_; }
    inline void Render(const DepthShader &s) const;
    inline void Render(const GShader &s) const;
    inline void Render(const BlinnPhongShader &s) const;

private:
    int id_ = -1;
    bool pbr_ = false;
};

// ---------------- render targets (src/buffer.h) ----------------

// A host copy of render-target planes.
struct Buffer {
    int w = 0, h = 0, channels = 1;
    std::vector<Float> data;   // plane-major: channel c of pixel (x, y) at c*w*h + x + y*w
    int GetWidth() const { return w; }
    int GetHeight() const { return h; }
    Float GetValue(int x, int y) const { return data[(size_t)x + (size_t)y * w]; }
    Vector3f GetValue3(int x, int y) const
    {
        const size_t P = (size_t)w * h, i = (size_t)x + (size_t)y * w;
        return Vector3f(data[i], data[P + i], data[2 * P + i]);
    }
};

// One of ForkerGL's named buffers, backed by a device target: plane0 and
// `planes` select the buffer's planes inside the target (the seven G-buffers
// are planes of one 19-plane target).
struct TargetBuffer {
    int which, plane0, planes;
    inline Buffer Read() const;
    Buffer operator()() const { return Read(); }
    inline int GetWidth() const;
    inline int GetHeight() const;
};
struct TargetBuffer1f : TargetBuffer {
    Float GetValue(int x, int y) const { return Read().GetValue(x, y); }
};
struct TargetBuffer3f : TargetBuffer {
    Vector3f GetValue(int x, int y) const { return Read().GetValue3(x, y); }
};

// ---------------- ForkerGL (src/forkergl.h:15-81) ----------------

struct ForkerGL {
    ForkerGL() = delete;
    enum RenderMode { Forward, Deferred };
    enum PassType { ForwardPass, GeometryPass, LightingPass, ShadowPass };

    static fgl_ctx *Context(int device = 0)
    {
        static fgl_ctx *ctx = nullptr;
        if (!ctx && fgl_ctx_create(device, &ctx) != FGL_OK) throw std::runtime_error("fgl_ctx_create failed");
        return ctx;
    }

    // Texture modes captured by the textures of scenes loaded afterwards
    // (src/forkergl.cpp:15-16, 44-52; src/model.cpp:425-426).
    inline static Texture::WrapMode TextureWrapping = Texture::NoWrap;
    inline static Texture::FilterMode TextureFiltering = Texture::Nearest;
    static void TextureWrapMode(Texture::WrapMode m) { TextureWrapping = m; }
    static void TextureFilterMode(Texture::FilterMode m) { TextureFiltering = m; }

    // The reference's public buffers (src/forkergl.h:41-54).
    inline static const TargetBuffer3f FrameBuffer{{FGL_TARGET_FRAME, 0, 3}};
    inline static const TargetBuffer1f DepthBuffer{{FGL_TARGET_DEPTH, 0, 1}};
    inline static const TargetBuffer1f ShadowBuffer{{FGL_TARGET_SHADOW, 0, 1}};
    inline static const TargetBuffer3f NormalGBuffer{{FGL_TARGET_GBUFFER, 0, 3}};
    inline static const TargetBuffer3f WorldPosGBuffer{{FGL_TARGET_GBUFFER, 3, 3}};
    inline static const TargetBuffer3f LightSpaceNDCPosGBuffer{{FGL_TARGET_GBUFFER, 6, 3}};
    inline static const TargetBuffer3f AlbedoGBuffer{{FGL_TARGET_GBUFFER, 9, 3}};
    inline static const TargetBuffer3f EmissiveGBuffer{{FGL_TARGET_GBUFFER, 12, 3}};
    inline static const TargetBuffer3f ParamGBuffer{{FGL_TARGET_GBUFFER, 15, 3}};
    inline static const TargetBuffer1f ShadingTypeGBuffer{{FGL_TARGET_GBUFFER, 18, 1}};
    inline static const TargetBuffer1f AmbientOcclusionGBuffer{{FGL_TARGET_AO, 0, 1}};
    // The 19 G-buffer planes at once (normal, world position, light-space NDC,
    // albedo, emissive, param: 3 each; shading type: 1).
    inline static const TargetBuffer GeometryBuffers{FGL_TARGET_GBUFFER, 0, 19};
    // AntiAliasedImage (src/render.cpp:291-343): 8-bit B,G,R after DoSSAA.
    static std::vector<uint8_t> AntiAliasedImage()
    {
        size_t n = 0;
        check(Context(), fgl_target_bytes(Context(), FGL_TARGET_IMAGE, &n), "AntiAliasedImage");
        std::vector<uint8_t> img(n);
        check(Context(), fgl_read_target(Context(), FGL_TARGET_IMAGE, img.data(), n), "AntiAliasedImage");
        return img;
    }

    // sizes of the frame targets and of the shadow map, as last initialised
    inline static int Width = 0, Height = 0, ShadowWidth = 0, ShadowHeight = 0;

    static void InitFrameBuffer(int w, int h) { Init(FGL_TARGET_FRAME, w, h, "InitFrameBuffer"); }
    static void InitDepthBuffer(int w, int h) { Init(FGL_TARGET_DEPTH, w, h, "InitDepthBuffer"); }
    static void InitGeometryBuffers(int w, int h) { Init(FGL_TARGET_GBUFFER, w, h, "InitGeometryBuffers"); }
    static void InitShadowBuffer(int w, int h)
    {
        ShadowWidth = w;
        ShadowHeight = h;
        check(Context(), fgl_init_shadow_buffer(Context(), w, h), "InitShadowBuffer");
    }
    static void ClearColor(const Color3 &c) { check(Context(), fgl_clear_color(Context(), c.x, c.y, c.z), "ClearColor"); }

    // SetViewportMatrix (src/forkergl.cpp:89-102): the shadow pass shares the
    // frame's viewport in the reference (its shadow map is frame-sized).
    static void SetViewportMatrix(int x, int y, int w, int h)
    {
        check(Context(), fgl_set_viewport(Context(), 0, x, y, w, h), "SetViewportMatrix");
        check(Context(), fgl_set_viewport(Context(), 1, x, y, w, h), "SetViewportMatrix");
        Matrix4x4f m;
        m(0, 0) = w / 2.f; m(0, 3) = x + w / 2.f;
        m(1, 1) = h / 2.f; m(1, 3) = y + h / 2.f;
        m(2, 2) = 1 / 2.f; m(2, 3) = 1 / 2.f;
        viewport_ = m;
    }
    static Matrix4x4f GetViewportMatrix() { return viewport_; }
    // Extension (the `shadowmap S` scene key): a shadow map of its own size.
    static void SetShadowViewportMatrix(int x, int y, int w, int h)
    {
        check(Context(), fgl_set_viewport(Context(), 1, x, y, w, h), "SetShadowViewportMatrix");
    }
    static void SetViewProjectionMatrix(const Matrix4x4f &m) { view_projection_ = m; }
    static Matrix4x4f GetViewProjectionMatrix() { return view_projection_; }
    static void SetLightSpaceMatrix(const Matrix4x4f &m)
    {
        light_space_ = m;
        check(Context(), fgl_set_matrix(Context(), FGL_MAT_LIGHT_SPACE, m.data()), "SetLightSpaceMatrix");
    }
    static Matrix4x4f GetLightSpaceMatrix() { return light_space_; }
    static void SetRenderMode(RenderMode m) { mode_ = m; }
    static RenderMode GetRenderMode() { return mode_; }
    static void SetPassType(PassType t)
    {
        static const int map[] = {FGL_PASS_FORWARD, FGL_PASS_GEOMETRY, FGL_PASS_LIGHTING, FGL_PASS_SHADOW};
        check(Context(), fgl_begin_pass(Context(), map[t]), "SetPassType");
    }
    // DrawScreenSpacePixels (src/forkergl.cpp:326-380): the scene's camera
    // position and lights shade the G-buffer.
    static inline void DrawScreenSpacePixels(const Scene &scene);

private:
    static void Init(int which, int w, int h, const char *what)
    {
        Width = w;
        Height = h;
        check(Context(), fgl_init_buffer(Context(), which, w, h), what);
    }
    inline static Matrix4x4f viewport_, view_projection_, light_space_;
    inline static RenderMode mode_ = Forward;
};

// Shadow::SetShadowStatus / GetShadowStatus and the filter parameters of
// src/shaders/shadow.h:12-23 (src/shaders/shadow.cpp:9-21).  The filter is
// the reference's default build's until a Scene or a caller sets another:
// SOFT_SHADOW_PCSS with its random disk samples (32 blocker-search samples
// over 0.005, 64 PCF samples over 0.007 times the penumbra, area light 2.5).
struct Shadow {
    static fgl_shadow_cfg ReferencePCSS()
    {
        fgl_shadow_cfg c{};
        c.mode = FGL_SHADOW_REF_PCSS;
        c.n_blocker = 32;
        c.n_pcf = 64;
        c.pcf_filter = 0.007f;
        c.blocker_filter = 0.005;
        c.pcss_filter = 0.007;
        c.area_light = 2.5f;
        return c;
    }
    static void SetShadowStatus(bool on) { SetShadowStatus(on, cfg_); }
    static void SetShadowStatus(bool on, const fgl_shadow_cfg &cfg)
    {
        on_ = on;
        cfg_ = cfg;
        check(ForkerGL::Context(), fgl_set_shadow(ForkerGL::Context(), on ? 1 : 0, &cfg), "SetShadowStatus");
    }
    static bool GetShadowStatus() { return on_; }

private:
    inline static bool on_ = true;
    inline static fgl_shadow_cfg cfg_ = ReferencePCSS();
};

// ---------------- Scene (src/scene.h) ----------------

// A .scene file loaded with the texture modes ForkerGL holds at that moment
// and uploaded into the facade's context.  Loading applies the scene's
// render mode and shadow settings, as the reference's Scene constructor
// does (src/scene.cpp:54-196).
class Scene {
public:
    explicit Scene(const std::string &path)
    {
        char err[512] = {0};
        if (fgl_scene_load_with(path.c_str(), ForkerGL::TextureWrapping, ForkerGL::TextureFiltering, &s_, err,
                                sizeof(err)) != FGL_OK)
            throw std::runtime_error(std::string("Scene: ") + err);
        fgl_scene_info_get(s_, &info_);
        std::vector<int32_t> ids(info_.n_models);
        check(ForkerGL::Context(), fgl_scene_upload(ForkerGL::Context(), s_, ids.data()), "Scene upload");
        for (int i = 0; i < info_.n_models; ++i) {
            fgl_model_desc d;
            Matrix4x4f M;
            fgl_scene_model(s_, i, &d, M.data());
            models_.emplace_back(ids[i], d.support_pbr != 0);
            matrices_.push_back(M);
        }
        camera_ = Camera(Point3f(info_.eye[0], info_.eye[1], info_.eye[2]),
                         Point3f(info_.look_at[0], info_.look_at[1], info_.look_at[2]));
        for (int j = 0; j < 3 && info_.n_lights > 0; ++j) {
            light_.position[j] = info_.lights[0].vec[j];
            light_.color[j] = info_.lights[0].color[j];
        }
        ForkerGL::SetRenderMode(info_.deferred ? ForkerGL::Deferred : ForkerGL::Forward);
        Shadow::SetShadowStatus(info_.shadow_on != 0, info_.shadow);
    }
    ~Scene() { fgl_scene_free(s_); }
    Scene(const Scene &) = delete;
    Scene &operator=(const Scene &) = delete;

    int GetWidth() const { return info_.width; }
    int GetHeight() const { return info_.height; }
    Float GetRatio() const { return Float(info_.width) / info_.height; }
    bool IsSSAAOn() const { return info_.ssaa_on != 0; }
    int GetSSAAKernelSize() const { return info_.ssaa_k; }
    bool IsSSAOOn() const { return info_.ssao_on != 0; }
    const PointLight &GetPointLight() const { return light_; }
    const Camera &GetCamera() const { return camera_; }
    Camera::ProjectionType GetProjectionType() const
    {
        return info_.camera_ortho ? Camera::Orthographic : Camera::Perspective;
    }
    unsigned int GetModelCount() const { return (unsigned int)models_.size(); }
    const Model &GetModel(int i) const { return models_[i]; }
    Matrix4x4f GetModelMatrix(int i) const { return matrices_[i]; }
    // the loaded file's full settings, extensions included (fgl_scene_info)
    const fgl_scene_info &Info() const { return info_; }

private:
    fgl_scene *s_ = nullptr;
    fgl_scene_info info_{};
    std::vector<Model> models_;
    std::vector<Matrix4x4f> matrices_;
    Camera camera_;
    PointLight light_;
};

// Render::DoSSAO + the AO blur (src/render.cpp:214-289) and Render::DoSSAA
// (src/render.cpp:291-343) for a loaded scene: the frame driver's two
// screen-space steps with the scene's SSAO settings and SSAA factor.
inline void DoSSAO(const Scene &scene)
{
    check(ForkerGL::Context(), fgl_ssao(ForkerGL::Context(), &scene.Info().ssao), "DoSSAO");
}
inline void DoSSAA(const Scene &scene)
{
    const int k = scene.IsSSAAOn() ? scene.GetSSAAKernelSize() : 1;
    check(ForkerGL::Context(), fgl_resolve(ForkerGL::Context(), k), "DoSSAA");
}

// ---------------- definitions that need the types above ----------------

inline void Model::Render(const DepthShader &s) const
{
    fgl_ctx *c = ForkerGL::Context();
    check(c, fgl_set_matrix(c, FGL_MAT_LIGHT_SPACE, s.uLightSpaceMatrix.data()), "uLightSpaceMatrix");
    check(c, fgl_draw_model(c, id_, s.uModelMatrix.data()), "Model::Render");
}

inline void Model::Render(const GShader &s) const
{
    fgl_ctx *c = ForkerGL::Context();
    check(c, fgl_set_matrix(c, FGL_MAT_VIEW, s.uViewMatrix.data()), "uViewMatrix");
    check(c, fgl_set_matrix(c, FGL_MAT_PROJ, s.uProjectionMatrix.data()), "uProjectionMatrix");
    check(c, fgl_set_matrix(c, FGL_MAT_LIGHT_SPACE, s.uLightSpaceMatrix.data()), "uLightSpaceMatrix");
    check(c, fgl_draw_model_normal(c, id_, s.uModelMatrix.data(), s.uNormalMatrix.data()), "Model::Render");
}

inline void Model::Render(const BlinnPhongShader &s) const
{
    fgl_ctx *c = ForkerGL::Context();
    check(c, fgl_set_matrix(c, FGL_MAT_VIEW, s.uViewMatrix.data()), "uViewMatrix");
    check(c, fgl_set_matrix(c, FGL_MAT_PROJ, s.uProjectionMatrix.data()), "uProjectionMatrix");
    check(c, fgl_set_matrix(c, FGL_MAT_LIGHT_SPACE, s.uLightSpaceMatrix.data()), "uLightSpaceMatrix");
    check(c, fgl_set_eye(c, s.uEyePos.data()), "uEyePos");
    fgl_light l{};
    l.type = 0;
    std::memcpy(l.vec, s.uPointLight.position.data(), sizeof(l.vec));
    std::memcpy(l.color, s.uPointLight.color.data(), sizeof(l.color));
    check(c, fgl_set_lights(c, 1, &l), "uPointLight");
    check(c, fgl_draw_model_normal(c, id_, s.uModelMatrix.data(), s.uNormalMatrix.data()), "Model::Render");
}

inline void ForkerGL::DrawScreenSpacePixels(const Scene &scene)
{
    fgl_ctx *c = Context();
    const Point3f eye = scene.GetCamera().GetPosition();
    check(c, fgl_set_eye(c, eye.data()), "DrawScreenSpacePixels");
    const fgl_scene_info &I = scene.Info();
    check(c, fgl_set_lights(c, I.n_lights, I.lights), "DrawScreenSpacePixels");
    check(c, fgl_draw_screen_space_pixels(c), "DrawScreenSpacePixels");
}

inline Buffer TargetBuffer::Read() const
{
    fgl_ctx *c = ForkerGL::Context();
    size_t n = 0;
    check(c, fgl_target_bytes(c, which, &n), "read");
    std::vector<Float> all(n / sizeof(Float));
    check(c, fgl_read_target(c, which, all.data(), n), "read");
    Buffer b;
    b.w = GetWidth();
    b.h = GetHeight();
    b.channels = planes;
    const size_t P = (size_t)b.w * b.h;
    b.data.assign(all.begin() + plane0 * P, all.begin() + (plane0 + planes) * P);
    return b;
}
inline int TargetBuffer::GetWidth() const { return which == FGL_TARGET_SHADOW ? ForkerGL::ShadowWidth : ForkerGL::Width; }
inline int TargetBuffer::GetHeight() const
{
    return which == FGL_TARGET_SHADOW ? ForkerGL::ShadowHeight : ForkerGL::Height;
}

}  // namespace fglx
