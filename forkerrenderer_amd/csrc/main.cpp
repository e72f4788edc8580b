// main.cpp -- ForkerRendererAMD <scene file>: the reference's command line
// (src/main.cpp:18-55) on the GPU path.  Loads the scene, renders it
// (Render::Render), and writes the reference's output images
// (src/output.cpp:12-86) into ./output/.
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "fgl.h"

// TGAColor(v) from a float: a truncating 32-bit conversion, low byte kept
// (src/buffer.cpp:20-32, 113-126).
static uint8_t to_byte(float v)
{
    int i = (v >= -2147483648.0f && v < 2147483648.0f) ? (int)v : (int)0x80000000;
    return (uint8_t)(i & 255);
}

static bool read_target(fgl_ctx *c, int which, std::vector<float> &out)
{
    size_t n = 0;
    if (fgl_target_bytes(c, which, &n) != FGL_OK || n == 0) return false;
    out.resize(n / 4);
    return fgl_read_target(c, which, out.data(), n) == FGL_OK;
}

static void write_rgb(const char *path, const float *planes, int w, int h)
{
    size_t P = (size_t)w * h;
    std::vector<uint8_t> img(P * 3);
    for (size_t i = 0; i < P; ++i) {
        img[3 * i + 0] = to_byte(planes[2 * P + i] * 254.99f);
        img[3 * i + 1] = to_byte(planes[P + i] * 254.99f);
        img[3 * i + 2] = to_byte(planes[i] * 254.99f);
    }
    fgl_tga_write(path, w, h, 3, img.data(), 1);
    std::printf("[info] Output TGA File: '%s'\n", path);
}

static void write_grey(const char *path, const float *plane, int w, int h)
{
    size_t P = (size_t)w * h;
    std::vector<uint8_t> img(P);
    for (size_t i = 0; i < P; ++i) img[i] = to_byte(plane[i] * 255);
    fgl_tga_write(path, w, h, 1, img.data(), 1);
    std::printf("[info] Output TGA File: '%s'\n", path);
}

int main(int argc, const char *argv[])
{
    if (argc != 2) {
        std::fprintf(stderr, "Required: 1 argument, but given %d\nUsage: ./ForkerRendererAMD <scene file>\n",
                     argc - 1);
        return 1;
    }
    auto t0 = std::chrono::steady_clock::now();
    fgl_scene *scene = nullptr;
    char err[1024] = {0};
    if (fgl_scene_load(argv[1], &scene, err, sizeof(err)) != FGL_OK) {
        std::fprintf(stderr, "[error] %s\n", err);
        return 1;
    }
    fgl_scene_info I;
    fgl_scene_info_get(scene, &I);
    auto t1 = std::chrono::steady_clock::now();
    std::printf("[info] Scene File: '%s' (%d x %d, %s, SSAA %s x%d, shadow %s, SSAO %s, %d models)\n", argv[1],
                I.width, I.height, I.deferred ? "deferred" : "forward", I.ssaa_on ? "on" : "off", I.ssaa_k,
                I.shadow_on ? "on" : "off", I.ssao_on ? "on" : "off", I.n_models);
    std::printf("[info] Time Used: %.6f Seconds (Scene Loaded)\n", std::chrono::duration<double>(t1 - t0).count());
    fgl_ctx *ctx = nullptr;
    if (fgl_ctx_create(0, &ctx) != FGL_OK) {
        std::fprintf(stderr, "[error] no GPU device\n");
        return 1;
    }
    auto t2 = std::chrono::steady_clock::now();
    if (fgl_render_scene(ctx, scene) != FGL_OK || fgl_sync(ctx) != FGL_OK) {
        std::fprintf(stderr, "[error] render failed: %s\n", fgl_last_error(ctx));
        return 1;
    }
    auto t3 = std::chrono::steady_clock::now();
    std::printf("[info] Time Used: %.6f Seconds (Render, including upload)\n",
                std::chrono::duration<double>(t3 - t2).count());
    mkdir("output", 0755);
    const int k = I.ssaa_on ? I.ssaa_k : 1;
    const int W = I.width * k, H = I.height * k;
    std::vector<float> buf;
    if (read_target(ctx, FGL_TARGET_FRAME, buf)) write_rgb("output/framebuffer.tga", buf.data(), W, H);
    if (I.ssaa_on) {
        size_t n = 0;
        fgl_target_bytes(ctx, FGL_TARGET_IMAGE, &n);
        std::vector<uint8_t> img(n);
        if (n && fgl_read_target(ctx, FGL_TARGET_IMAGE, img.data(), n) == FGL_OK) {
            fgl_tga_write("output/framebuffer_SSAA.tga", W / k, H / k, 3, img.data(), 1);
            std::printf("[info] Output TGA File: 'output/framebuffer_SSAA.tga'\n");
        }
    }
    if (I.shadow_on && read_target(ctx, FGL_TARGET_SHADOW, buf)) {
        int S = I.shadow_map_size > 0 ? I.shadow_map_size : 0;
        write_grey("output/shadowmap.tga", buf.data(), S ? S : W, S ? S : H);
    }
    if (read_target(ctx, FGL_TARGET_DEPTH, buf)) write_grey("output/zbuffer.tga", buf.data(), W, H);
    if (I.deferred && read_target(ctx, FGL_TARGET_GBUFFER, buf)) {
        size_t P = (size_t)W * H;
        write_rgb("output/gbuffer_normal.tga", buf.data(), W, H);
        write_rgb("output/gbuffer_worldpos.tga", buf.data() + 3 * P, W, H);
        write_rgb("output/gbuffer_albedo.tga", buf.data() + 9 * P, W, H);
        write_rgb("output/gbuffer_param.tga", buf.data() + 15 * P, W, H);
        write_grey("output/gbuffer_shading_type.tga", buf.data() + 18 * P, W, H);
        if (read_target(ctx, FGL_TARGET_AO, buf))
            write_grey("output/gbuffer_ambient_occlusion.tga", buf.data(), W, H);
    }
    fgl_ctx_destroy(ctx);
    fgl_scene_free(scene);
    return 0;
}
