// forkergl_amd.hpp -- C++ host facade with the reference's names over the C
// ABI in fgl.h, so code written against src/forkergl.h, src/shaders/*.h and
// Model::Render (src/render.cpp) ports by changing includes.
//
//   reference                                  facade
//   ForkerGL::InitFrameBuffer(w, h)            fglx::ForkerGL::InitFrameBuffer(w, h)
//   ForkerGL::SetViewportMatrix(x, y, w, h)    fglx::ForkerGL::SetViewportMatrix(x, y, w, h)
//   ForkerGL::SetPassType(ShadowPass)          fglx::ForkerGL::SetPassType(fglx::ForkerGL::ShadowPass)
//   DepthShader s; s.uModelMatrix = M; ...     fglx::DepthShader s; s.uModelMatrix = M; ...
//   model.Render(s)                            model.Render(s)   (model: fglx::Model)
//   ForkerGL::DrawScreenSpacePixels(scene)     fglx::ForkerGL::DrawScreenSpacePixels()
//   ForkerGL::FrameBuffer.GetValue(x, y)       fglx::ForkerGL::FrameBuffer().GetValue(x, y)
//
// Matrices are row-major float[16] (the reference's Matrix4x4f rows).  The
// per-face DrawTriangle of the reference has no facade: faces are submitted
// per model, in the reference's order, by Model::Render.
#pragma once

#include <array>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "fgl.h"

namespace fglx {

using Mat4 = std::array<float, 16>;
using Vec3 = std::array<float, 3>;

inline Mat4 Identity()
{
    return {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
}

inline void check(fgl_ctx *c, int rc, const char *what)
{
    if (rc != FGL_OK) throw std::runtime_error(std::string(what) + ": " + fgl_last_error(c));
}

// A host copy of one render target with the reference's accessors.
struct Buffer {
    int w = 0, h = 0, channels = 1;
    std::vector<float> data;  // plane-major (SoA)
    int GetWidth() const { return w; }
    int GetHeight() const { return h; }
    float GetValue(int x, int y) const { return data[(size_t)x + (size_t)y * w]; }
    Vec3 GetValue3(int x, int y) const
    {
        size_t P = (size_t)w * h, i = (size_t)x + (size_t)y * w;
        return {data[i], data[P + i], data[2 * P + i]};
    }
};

struct ForkerGL {
    enum RenderMode { Forward, Deferred };
    enum PassType { ForwardPass, GeometryPass, LightingPass, ShadowPass };

    static fgl_ctx *Context(int device = 0)
    {
        static fgl_ctx *ctx = nullptr;
        if (!ctx) {
            if (fgl_ctx_create(device, &ctx) != FGL_OK) throw std::runtime_error("fgl_ctx_create failed");
        }
        return ctx;
    }
    static RenderMode &Mode() { static RenderMode m = Forward; return m; }
    static int &Width() { static int w = 0; return w; }
    static int &Height() { static int h = 0; return h; }
    static Mat4 &Viewport() { static Mat4 m = Identity(); return m; }
    static Mat4 &ViewProjection() { static Mat4 m = Identity(); return m; }
    static Mat4 &LightSpace() { static Mat4 m = Identity(); return m; }

    static void InitFrameBuffer(int w, int h) { Width() = w; Height() = h; check(Context(), fgl_init_buffer(Context(), FGL_TARGET_FRAME, w, h), "InitFrameBuffer"); }
    static void InitDepthBuffer(int w, int h) { Width() = w; Height() = h; check(Context(), fgl_init_buffer(Context(), FGL_TARGET_DEPTH, w, h), "InitDepthBuffer"); }
    static void InitGeometryBuffers(int w, int h) { Width() = w; Height() = h; check(Context(), fgl_init_buffer(Context(), FGL_TARGET_GBUFFER, w, h), "InitGeometryBuffers"); }
    static void InitShadowBuffer(int w, int h) { check(Context(), fgl_init_shadow_buffer(Context(), w, h), "InitShadowBuffer"); }
    static void ClearColor(const Vec3 &c) { check(Context(), fgl_clear_color(Context(), c[0], c[1], c[2]), "ClearColor"); }
    static void SetViewportMatrix(int x, int y, int w, int h)
    {
        check(Context(), fgl_set_viewport(Context(), 0, x, y, w, h), "SetViewportMatrix");
        check(Context(), fgl_set_viewport(Context(), 1, x, y, w, h), "SetViewportMatrix");
        Mat4 m = Identity();
        m[0] = w / 2.f; m[5] = h / 2.f; m[3] = x + w / 2.f; m[7] = y + h / 2.f; m[10] = 1 / 2.f; m[11] = 1 / 2.f;
        Viewport() = m;
    }
    static Mat4 GetViewportMatrix() { return Viewport(); }
    // Extension (the `shadowmap S` scene key): a shadow map of its own size.
    static void SetShadowViewportMatrix(int x, int y, int w, int h)
    {
        check(Context(), fgl_set_viewport(Context(), 1, x, y, w, h), "SetShadowViewportMatrix");
    }
    static void SetViewProjectionMatrix(const Mat4 &m) { ViewProjection() = m; }
    static Mat4 GetViewProjectionMatrix() { return ViewProjection(); }
    static void SetLightSpaceMatrix(const Mat4 &m)
    {
        LightSpace() = m;
        check(Context(), fgl_set_matrix(Context(), FGL_MAT_LIGHT_SPACE, m.data()), "SetLightSpaceMatrix");
    }
    static Mat4 GetLightSpaceMatrix() { return LightSpace(); }
    static void SetRenderMode(RenderMode m) { Mode() = m; }
    static RenderMode GetRenderMode() { return Mode(); }
    static void SetPassType(PassType t)
    {
        static const int map[] = {FGL_PASS_FORWARD, FGL_PASS_GEOMETRY, FGL_PASS_LIGHTING, FGL_PASS_SHADOW};
        check(Context(), fgl_begin_pass(Context(), map[t]), "SetPassType");
    }
    static void DrawScreenSpacePixels() { check(Context(), fgl_draw_screen_space_pixels(Context()), "DrawScreenSpacePixels"); }

    static Buffer Read(int which, int channels)
    {
        Buffer b;
        size_t n = 0;
        check(Context(), fgl_target_bytes(Context(), which, &n), "target");
        b.data.resize(n / 4);
        check(Context(), fgl_read_target(Context(), which, b.data.data(), n), "read");
        b.channels = channels;
        b.w = Width();
        b.h = Height();
        return b;
    }
    static Buffer FrameBuffer() { return Read(FGL_TARGET_FRAME, 3); }
    static Buffer DepthBuffer() { return Read(FGL_TARGET_DEPTH, 1); }
    static Buffer AmbientOcclusionGBuffer() { return Read(FGL_TARGET_AO, 1); }
    // The seven G-buffers as 19 planes (normal, world position, light-space
    // NDC, albedo, emissive, param: 3 each; shading type: 1).
    static Buffer GeometryBuffers() { return Read(FGL_TARGET_GBUFFER, 19); }
    // AntiAliasedImage (render.cpp:291-343): 8-bit B,G,R after DoSSAA.
    static std::vector<uint8_t> AntiAliasedImage()
    {
        size_t n = 0;
        check(Context(), fgl_target_bytes(Context(), FGL_TARGET_IMAGE, &n), "target");
        std::vector<uint8_t> img(n);
        check(Context(), fgl_read_target(Context(), FGL_TARGET_IMAGE, img.data(), n), "read");
        return img;
    }
};

// Shadow::SetShadowStatus and the shadow.h filter parameters
// (src/shaders/shadow.cpp:11-21, shadow.h:12-23).
struct Shadow {
    static void SetShadowStatus(bool on, const fgl_shadow_cfg &cfg)
    {
        check(ForkerGL::Context(), fgl_set_shadow(ForkerGL::Context(), on ? 1 : 0, &cfg), "SetShadowStatus");
    }
};

// Render::DoSSAO + the AO blur (src/render.cpp:214-289) and Render::DoSSAA
// (src/render.cpp:291-343): the frame driver's two screen-space steps.
inline void DoSSAO(const fgl_ssao_cfg &cfg) { check(ForkerGL::Context(), fgl_ssao(ForkerGL::Context(), &cfg), "DoSSAO"); }
inline void DoSSAA(int k) { check(ForkerGL::Context(), fgl_resolve(ForkerGL::Context(), k), "DoSSAA"); }

// Shader uniform blocks (src/shaders/*.h).  The fragment programs run on the
// GPU; the struct fields keep the reference's names.
struct PointLight {
    Vec3 position{0, 0, 0};
    Vec3 color{1, 1, 1};
};
struct DepthShader {
    Mat4 uModelMatrix = Identity(), uLightSpaceMatrix = Identity();
};
struct GShader {
    Mat4 uModelMatrix = Identity(), uViewMatrix = Identity(), uProjectionMatrix = Identity(),
         uLightSpaceMatrix = Identity();
};
struct BlinnPhongShader {
    Mat4 uModelMatrix = Identity(), uViewMatrix = Identity(), uProjectionMatrix = Identity(),
         uLightSpaceMatrix = Identity();
    PointLight uPointLight;
    Vec3 uEyePos{0, 0, 0};
};
using PBRShader = BlinnPhongShader;

// A model uploaded to the context (Model::Load's product).
struct Model {
    int id = -1;
    void Render(const DepthShader &s) const { Draw(s.uModelMatrix); }
    void Render(const GShader &s) const
    {
        fgl_ctx *c = ForkerGL::Context();
        check(c, fgl_set_matrix(c, FGL_MAT_VIEW, s.uViewMatrix.data()), "uViewMatrix");
        check(c, fgl_set_matrix(c, FGL_MAT_PROJ, s.uProjectionMatrix.data()), "uProjectionMatrix");
        check(c, fgl_set_matrix(c, FGL_MAT_LIGHT_SPACE, s.uLightSpaceMatrix.data()), "uLightSpaceMatrix");
        Draw(s.uModelMatrix);
    }
    void Render(const BlinnPhongShader &s) const
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
        Draw(s.uModelMatrix);
    }
    void Draw(const Mat4 &m) const
    {
        fgl_ctx *c = ForkerGL::Context();
        check(c, fgl_draw_model(c, id, m.data()), "Model::Render");
    }
};

}  // namespace fglx
