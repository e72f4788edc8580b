// facade_render.cpp -- facade_render <scene file> <output dir>: the
// reference's frame driver Render::Render (src/render.cpp:40-58, 60-212,
// 291-343) written against the C++ facade (include/forkergl_amd.hpp) the way
// a maintainer ports the reference's pass code: ForkerGL statics, shader
// uniform structs and Model::Render.  The scene's arrays are uploaded through
// the C ABI (the facade has no loader: the reference's Model::Load runs on
// the host).  Writes the frame's targets as raw files for comparison:
// depth.f32, fb.f32 (3 planes), image.u8, and in deferred mode gbuf.f32 (19
// planes) and ao.f32.
#include <cstdio>
#include <string>
#include <vector>

#include "forkergl_amd.hpp"

using namespace fglx;

static Mat4 mat(const float *m)
{
    Mat4 r;
    for (int i = 0; i < 16; ++i) r[i] = m[i];
    return r;
}

static Mat4 mul(const Mat4 &a, const Mat4 &b)
{
    Mat4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            float s = 0.f;
            for (int k = 0; k < 4; ++k) s += a[4 * i + k] * b[4 * k + j];
            r[4 * i + j] = s;
        }
    return r;
}

static void dump(const std::string &path, const void *p, size_t bytes)
{
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f || std::fwrite(p, 1, bytes, f) != bytes) throw std::runtime_error("cannot write " + path);
    std::fclose(f);
}

int main(int argc, const char *argv[])
{
    if (argc != 3) {
        std::fprintf(stderr, "Usage: facade_render <scene file> <output dir>\n");
        return 1;
    }
    try {
        fgl_scene *scene = nullptr;
        char err[512] = {0};
        if (fgl_scene_load(argv[1], &scene, err, sizeof(err)) != FGL_OK) throw std::runtime_error(err);
        fgl_scene_info I;
        fgl_scene_info_get(scene, &I);
        if (I.n_lights != 1 || I.lights[0].type != 0) throw std::runtime_error("one point light expected");
        fgl_ctx *c = ForkerGL::Context();
        // Model::Load's products, in the scene's order (ids 0.. match the
        // scene's texture and material indices)
        for (int i = 0; i < I.n_textures; ++i) {
            int w, h, bpp, wrap, filter, id;
            const uint8_t *bytes;
            fgl_scene_texture(scene, i, &w, &h, &bpp, &wrap, &filter, &bytes);
            check(c, fgl_upload_texture(c, w, h, bpp, bytes, wrap, filter, &id), "texture");
        }
        for (int i = 0; i < I.n_materials; ++i) {
            fgl_material m;
            int id;
            fgl_scene_material(scene, i, &m);
            check(c, fgl_upload_material(c, &m, &id), "material");
        }
        std::vector<Model> models(I.n_models);
        std::vector<Mat4> modelMatrix(I.n_models);
        std::vector<bool> pbr(I.n_models);
        for (int i = 0; i < I.n_models; ++i) {
            fgl_model_desc d;
            float M[16];
            fgl_scene_model(scene, i, &d, M);
            check(c, fgl_upload_model(c, &d, &models[i].id), "model");
            modelMatrix[i] = mat(M);
            pbr[i] = d.support_pbr != 0;
        }
        const int k = I.ssaa_on ? I.ssaa_k : 1;
        const int W = I.width * k, H = I.height * k;
        const float ratio = (float)I.width / I.height;
        const int S = I.shadow_map_size > 0 ? I.shadow_map_size : 0;

        // Preconfigure
        ForkerGL::SetViewportMatrix(0, 0, W, H);
        if (S) ForkerGL::SetShadowViewportMatrix(0, 0, S, S);
        Shadow::SetShadowStatus(I.shadow_on != 0, I.shadow);
        ForkerGL::SetRenderMode(I.deferred ? ForkerGL::Deferred : ForkerGL::Forward);
        PointLight light;
        for (int j = 0; j < 3; ++j) {
            light.position[j] = I.lights[0].vec[j];
            light.color[j] = I.lights[0].color[j];
        }
        Vec3 eye{I.eye[0], I.eye[1], I.eye[2]};

        // DoShadowPass
        if (I.shadow_on) {
            ForkerGL::InitShadowBuffer(S ? S : W, S ? S : H);
            ForkerGL::InitDepthBuffer(W, H);
            ForkerGL::SetPassType(ForkerGL::ShadowPass);
            float V[16], P[16];
            const float origin[3] = {0.f, 0.f, 0.f}, up[3] = {0.f, 1.f, 0.f};
            fgl_make_lookat(I.lights[0].vec, origin, up, V);
            fgl_make_ortho(-3.0f * ratio, 3.0f * ratio, -3.0f, 3.0f, 0.1f, 20.f, P);
            ForkerGL::SetLightSpaceMatrix(mul(mat(P), mat(V)));
            for (int i = 0; i < I.n_models; ++i) {
                DepthShader depthShader;
                depthShader.uModelMatrix = modelMatrix[i];
                depthShader.uLightSpaceMatrix = ForkerGL::GetLightSpaceMatrix();
                models[i].Render(depthShader);
            }
        }

        // camera
        float Vv[16], Pv[16];
        const float up[3] = {0.f, 1.f, 0.f};
        fgl_make_lookat(I.eye, I.look_at, up, Vv);
        if (I.camera_ortho) fgl_make_ortho(-1.f * ratio, 1.f * ratio, -1.f, 1.f, 0.01f, 20.f, Pv);
        else fgl_make_perspective(45.f, ratio, 0.01f, 20.f, Pv);
        const Mat4 viewMatrix = mat(Vv), projectionMatrix = mat(Pv);

        if (ForkerGL::GetRenderMode() == ForkerGL::Forward) {
            // DoForwardPass
            ForkerGL::InitFrameBuffer(W, H);
            ForkerGL::InitDepthBuffer(W, H);
            ForkerGL::ClearColor({0.12f, 0.12f, 0.12f});
            ForkerGL::SetPassType(ForkerGL::ForwardPass);
            ForkerGL::SetViewProjectionMatrix(mul(projectionMatrix, viewMatrix));
            for (int i = 0; i < I.n_models; ++i) {
                BlinnPhongShader shader;   // PBRShader has the same uniforms
                shader.uModelMatrix = modelMatrix[i];
                shader.uViewMatrix = viewMatrix;
                shader.uProjectionMatrix = projectionMatrix;
                shader.uPointLight = light;
                shader.uEyePos = eye;
                if (I.shadow_on) shader.uLightSpaceMatrix = ForkerGL::GetLightSpaceMatrix();
                models[i].Render(shader);  // the model's material picks Blinn-Phong or PBR
                (void)pbr;
            }
        } else {
            // DoGeometryPass
            ForkerGL::InitGeometryBuffers(W, H);
            ForkerGL::InitDepthBuffer(W, H);
            ForkerGL::SetPassType(ForkerGL::GeometryPass);
            ForkerGL::SetViewProjectionMatrix(mul(projectionMatrix, viewMatrix));
            for (int i = 0; i < I.n_models; ++i) {
                GShader geometryShader;
                geometryShader.uModelMatrix = modelMatrix[i];
                geometryShader.uViewMatrix = viewMatrix;
                geometryShader.uProjectionMatrix = projectionMatrix;
                geometryShader.uLightSpaceMatrix = ForkerGL::GetLightSpaceMatrix();
                models[i].Render(geometryShader);
            }
            // DoLightingPass
            ForkerGL::InitFrameBuffer(W, H);
            ForkerGL::SetPassType(ForkerGL::LightingPass);
            check(c, fgl_set_eye(c, I.eye), "eye");
            fgl_light l = I.lights[0];
            check(c, fgl_set_lights(c, 1, &l), "lights");
            if (I.ssao_on) DoSSAO(I.ssao);
            ForkerGL::DrawScreenSpacePixels();
        }
        // DoSSAA (k = 1 quantises only)
        DoSSAA(k);

        const std::string out = argv[2];
        Buffer depth = ForkerGL::DepthBuffer();
        dump(out + "/depth.f32", depth.data.data(), depth.data.size() * 4);
        Buffer fb = ForkerGL::FrameBuffer();
        dump(out + "/fb.f32", fb.data.data(), fb.data.size() * 4);
        std::vector<uint8_t> img = ForkerGL::AntiAliasedImage();
        dump(out + "/image.u8", img.data(), img.size());
        if (I.deferred) {
            Buffer g = ForkerGL::GeometryBuffers();
            dump(out + "/gbuf.f32", g.data.data(), g.data.size() * 4);
            Buffer ao = ForkerGL::AmbientOcclusionGBuffer();
            dump(out + "/ao.f32", ao.data.data(), ao.data.size() * 4);
        }
        fgl_ctx_destroy(c);
        fgl_scene_free(scene);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "[error] %s\n", e.what());
        return 1;
    }
    return 0;
}
