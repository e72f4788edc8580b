// render_port_main.cpp -- main() of oracle/_ref/render_port (TEST
// INFRASTRUCTURE): loads a scene through the facade, runs the reference's
// own Render::Preconfigure and Render::Render (compiled from the reference's
// src/render.cpp against include/forkergl_amd.hpp by oracle/ref_render_port.py)
// and writes the frame's targets as raw files: depth.f32, fb.f32, image.u8,
// shadow.f32 when shadows are on, gbuf.f32 and ao.f32 in deferred mode.
#include <cstdio>
#include <string>
#include <vector>

#include "forkergl_amd.hpp"

namespace Render {
void Preconfigure(const fglx::Scene &scene);
void Render(const fglx::Scene &scene);
}

static void dump(const std::string &path, const void *p, size_t bytes)
{
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f || std::fwrite(p, 1, bytes, f) != bytes) throw std::runtime_error("cannot write " + path);
    std::fclose(f);
}

int main(int argc, const char *argv[])
{
    using namespace fglx;
    if (argc != 3) {
        std::fprintf(stderr, "Usage: render_port <scene file> <output dir>\n");
        return 1;
    }
    try {
        Scene scene(argv[1]);          // the reference's main: Scene, then Preconfigure, then Render
        Render::Preconfigure(scene);
        Render::Render(scene);
        const std::string out = argv[2];
        const Buffer depth = ForkerGL::DepthBuffer.Read(), fb = ForkerGL::FrameBuffer.Read();
        dump(out + "/depth.f32", depth.data.data(), depth.data.size() * 4);
        dump(out + "/fb.f32", fb.data.data(), fb.data.size() * 4);
        const std::vector<uint8_t> img = ForkerGL::AntiAliasedImage();
        dump(out + "/image.u8", img.data(), img.size());
        if (Shadow::GetShadowStatus()) {
            const Buffer s = ForkerGL::ShadowBuffer.Read();
            dump(out + "/shadow.f32", s.data.data(), s.data.size() * 4);
        }
        if (ForkerGL::GetRenderMode() == ForkerGL::Deferred) {
            const Buffer g = ForkerGL::GeometryBuffers.Read(), ao = ForkerGL::AmbientOcclusionGBuffer.Read();
            dump(out + "/gbuf.f32", g.data.data(), g.data.size() * 4);
            dump(out + "/ao.f32", ao.data.data(), ao.data.size() * 4);
        }
        fgl_ctx_destroy(ForkerGL::Context());
    } catch (const std::exception &e) {
        std::fprintf(stderr, "[error] %s\n", e.what());
        return 1;
    }
    return 0;
}
