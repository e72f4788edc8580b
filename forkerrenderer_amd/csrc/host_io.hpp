// host_io.hpp -- host-side scene data: .scene/.obj/.mtl parsing, TGA codec
// and the reference's host matrix builders.
#pragma once

#include <array>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/fgl.h"

namespace fgl {

struct HostTexture {
    int w = 0, h = 0, bpp = 0, wrap = 0, filter = 0;
    std::vector<uint8_t> bytes;  // B,G,R[,A], (x + y*w)*bpp, after load flips
};

struct HostModel {
    std::vector<float> pos, uv, nrm, tan;
    std::vector<int32_t> vi, ti, ni, face_mat;  // faces in draw order
    bool has_tangents = false, support_pbr = false;
    int n_faces() const { return (int)vi.size() / 3; }
};

struct HostScene {
    fgl_scene_info info{};
    int multilight = 0;
    std::vector<std::shared_ptr<HostModel>> models;  // per .scene model line
    std::vector<std::array<float, 16>> model_mats;
    std::vector<fgl_material> materials;
    std::vector<HostTexture> textures;
    // device ids after upload into a context
    void *uploaded_ctx = nullptr;
    std::vector<int> dev_models;
};

bool tga_read(const std::string &path, HostTexture &t, std::string &err);
bool tga_write(const std::string &path, int w, int h, int bpp, const uint8_t *bytes, bool rle);
// wrap / filter: the texture modes textures take unless a `texture` key sets
// them (the reference's ForkerGL statics at load time: NoWrap, Nearest).
bool load_scene(const std::string &path, HostScene &s, std::string &err, int wrap = 0, int filter = 0);

void make_model_matrix(float tx, float ty, float tz, float roty_deg, float scale, float *out);
void make_lookat(const float *eye, const float *center, const float *up, float *out);
void make_perspective(float fov, float aspect, float n, float f, float *out);
void make_ortho(float l, float r, float b, float t, float n, float f, float *out);
void make_normal_matrix(const float *m, float *out9);
void make_viewport(int x, int y, int w, int h, float *out);
void mat_mul(const float *a, const float *b, float *out);

}  // namespace fgl
