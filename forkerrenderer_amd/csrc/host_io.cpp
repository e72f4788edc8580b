// host_io.cpp -- .scene / .obj / .mtl / .tga reading, TGA writing, and the
// reference's host-side matrix builders, restated for the drop-in.
//
// The parsers follow the reference's line-oriented grammar and its
// submission order: models in .scene line order, meshes of a model in
// std::map order of their "g" names, faces in file order with fan
// triangulation (src/scene.cpp:19-196, src/model.cpp:16-476).
#include "host_io.hpp"

#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

#include "fgl_math.hpp"

namespace fgl {

// ---------------- matrices (src/geometry.cpp:60-179) ----------------

static void identity(float *m)
{
    for (int i = 0; i < 16; ++i) m[i] = 0.f;
    m[0] = m[5] = m[10] = m[15] = 1.f;
}

void mat_mul(const float *a, const float *b, float *out)
{
    float t[16];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            t[4 * i + j] = dot4(vec4(a[4 * i], a[4 * i + 1], a[4 * i + 2], a[4 * i + 3]),
                                vec4(b[j], b[4 + j], b[8 + j], b[12 + j]));
    std::memcpy(out, t, sizeof(t));
}

// Cofactor expansion along row 0 with the sign applied to the minor's
// determinant (Matrix::Det / Cofactor, src/geometry.h:580-627).
static float det_rec(const float *m, int n)
{
    if (n == 1) return m[0];
    float ret = 0;
    float minor[16];
    for (int i = 0; i < n; ++i) {
        int k = 0;
        for (int r = 1; r < n; ++r)
            for (int c = 0; c < n; ++c)
                if (c != i) minor[k++] = m[r * n + c];
        ret += m[i] * (det_rec(minor, n - 1) * (float)((i % 2) ? -1 : 1));
    }
    return ret;
}

void make_normal_matrix(const float *m, float *out9)
{
    float inv = (float)1 / det_rec(m, 4);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            float minor[9];
            int k = 0;
            for (int r = 0; r < 4; ++r)
                for (int c = 0; c < 4; ++c)
                    if (r != i && c != j) minor[k++] = m[r * 4 + c];
            out9[3 * i + j] = (det_rec(minor, 3) * (float)(((i + j) % 2) ? -1 : 1)) * inv;
        }
}

void make_model_matrix(float tx, float ty, float tz, float roty_deg, float scale, float *out)
{
    const float Pi = (float)3.14159265358979323846;
    float S[16], R[16], T[16], TR[16];
    identity(S);
    S[0] = S[5] = S[10] = scale;
    identity(R);
    float rad = roty_deg * Pi / 180.f;
    // unqualified cos/sin of a float resolve to the double C functions
    R[0] = (float)std::cos((double)rad);
    R[2] = (float)std::sin((double)rad);
    R[8] = (float)(-std::sin((double)rad));
    R[10] = (float)std::cos((double)rad);
    identity(T);
    T[3] = tx; T[7] = ty; T[11] = tz;
    mat_mul(T, R, TR);
    mat_mul(TR, S, out);
}

void make_lookat(const float *eye, const float *center, const float *up, float *out)
{
    Vec3 e = vec3(eye[0], eye[1], eye[2]), c = vec3(center[0], center[1], center[2]);
    Vec3 u = vec3(up[0], up[1], up[2]);
    Vec3 front = norm3(sub3(c, e));
    Vec3 right = norm3(cross3(front, u));
    Vec3 upv = norm3(cross3(right, front));
    float R[16], T[16];
    identity(R);
    R[0] = right.x; R[1] = right.y; R[2] = right.z; R[3] = 0.f;
    R[4] = upv.x; R[5] = upv.y; R[6] = upv.z; R[7] = 0.f;
    R[8] = -front.x; R[9] = -front.y; R[10] = -front.z; R[11] = 0.f;
    identity(T);
    T[3] = -e.x; T[7] = -e.y; T[11] = -e.z;
    mat_mul(R, T, out);
}

// Note: m[3][3] keeps the identity's 1 (the reference never clears it).
void make_perspective(float fov, float aspect, float n, float f, float *out)
{
    const float Pi = (float)3.14159265358979323846;
    float t = std::tan((fov / 2.f) * Pi / 180.f);
    identity(out);
    out[0] = 1.f / (aspect * t);
    out[5] = 1.f / t;
    out[10] = -(f + n) / (f - n);
    out[11] = -2 * f * n / (f - n);
    out[14] = -1;
}

void make_ortho(float l, float r, float b, float t, float n, float f, float *out)
{
    identity(out);
    out[0] = 2.f / (r - l);
    out[5] = 2.f / (t - b);
    out[10] = -2.f / (f - n);
    out[3] = -(r + l) / (r - l);
    out[7] = -(t + b) / (t - b);
    out[11] = -(f + n) / (f - n);
    out[15] = 1.f;
}

void make_viewport(int x, int y, int w, int h, float *out)
{
    identity(out);
    out[0] = w / 2.f;
    out[5] = h / 2.f;
    out[3] = x + w / 2.f;
    out[7] = y + h / 2.f;
    out[10] = 1 / 2.f;
    out[11] = 1 / 2.f;
}

// ---------------- TGA (src/tgaimage.cpp:43-365) ----------------

#pragma pack(push, 1)
struct TgaHeader {
    uint8_t idlength, colormaptype, datatypecode;
    uint16_t colormaporigin, colormaplength;
    uint8_t colormapdepth;
    uint16_t x_origin, y_origin, width, height;
    uint8_t bitsperpixel, imagedescriptor;
};
#pragma pack(pop)

static void flip_vertically(HostTexture &t)
{
    size_t line = (size_t)t.w * t.bpp;
    std::vector<uint8_t> tmp(line);
    for (int j = 0; j < t.h / 2; ++j) {
        uint8_t *a = t.bytes.data() + j * line, *b = t.bytes.data() + (size_t)(t.h - 1 - j) * line;
        std::memcpy(tmp.data(), a, line);
        std::memcpy(a, b, line);
        std::memcpy(b, tmp.data(), line);
    }
}

static void flip_horizontally(HostTexture &t)
{
    for (int i = 0; i < t.w / 2; ++i)
        for (int j = 0; j < t.h; ++j) {
            uint8_t *a = t.bytes.data() + ((size_t)i + (size_t)j * t.w) * t.bpp;
            uint8_t *b = t.bytes.data() + ((size_t)(t.w - 1 - i) + (size_t)j * t.w) * t.bpp;
            for (int c = 0; c < t.bpp; ++c) std::swap(a[c], b[c]);
        }
}

bool tga_read(const std::string &path, HostTexture &t, std::string &err)
{
    std::ifstream in(path, std::ios::binary);
    if (!in.is_open()) { err = "can't open file: " + path; return false; }
    TgaHeader h;
    in.read(reinterpret_cast<char *>(&h), sizeof(h));
    if (!in.good()) { err = "an error occurred while reading the header: " + path; return false; }
    t.w = h.width;
    t.h = h.height;
    t.bpp = h.bitsperpixel >> 3;
    if (t.w <= 0 || t.h <= 0 || (t.bpp != 1 && t.bpp != 3 && t.bpp != 4)) {
        err = "bad bpp (or width/height) value: " + path;
        return false;
    }
    size_t nbytes = (size_t)t.bpp * t.w * t.h;
    t.bytes.assign(nbytes, 0);
    if (h.datatypecode == 2 || h.datatypecode == 3) {
        in.read(reinterpret_cast<char *>(t.bytes.data()), nbytes);
        if (!in.good()) { err = "an error occurred while reading the data: " + path; return false; }
    } else if (h.datatypecode == 10 || h.datatypecode == 11) {
        // Packets: a header byte h < 128 is followed by h + 1 literal pixels;
        // h >= 128 by one pixel repeated h - 127 times.  A packet that would
        // put a pixel past the image is an error, reported before the write.
        const size_t npix = (size_t)t.w * t.h;
        size_t cur = 0;
        uint8_t px[4];
        while (cur < npix) {
            const int ch = in.get();
            if (!in.good()) { err = "an error occurred while reading the data: " + path; return false; }
            const bool literal = ch < 128;
            const int count = literal ? ch + 1 : ch - 127;
            for (int i = 0; i < count; ++i) {
                if (literal || i == 0) {
                    in.read(reinterpret_cast<char *>(px), t.bpp);
                    if (!in.good()) { err = "rle read error: " + path; return false; }
                }
                if (cur >= npix) { err = "too many pixels read: " + path; return false; }
                std::memcpy(t.bytes.data() + cur * t.bpp, px, t.bpp);
                ++cur;
            }
        }
    } else {
        err = "unknown file format " + std::to_string((int)h.datatypecode) + ": " + path;
        return false;
    }
    if (!(h.imagedescriptor & 0x20)) flip_vertically(t);
    if (h.imagedescriptor & 0x10) flip_horizontally(t);
    return true;
}

// Length of the run-length packet that starts at pixel i, as the reference's
// encoder chooses it (src/tgaimage.cpp:259-300; greedy, left to right, at
// most 128 pixels per packet).  When pixel i equals pixel i+1 the packet
// repeats pixel i over the whole stretch of equal pixels.  Otherwise it holds
// literal pixels and ends just before the first pixel at offset 1..126 that
// equals its successor (that pixel starts the next, repeating packet); with
// no such pixel it takes 128 pixels, or the rest of the image.
static size_t rle_packet(const uint8_t *d, size_t npix, int bpp, size_t i, bool &repeat)
{
    auto same = [&](size_t a, size_t b) { return std::memcmp(d + a * bpp, d + b * bpp, bpp) == 0; };
    const size_t left = npix - i, cap = left < 128 ? left : 128;
    size_t n = 1;
    repeat = cap >= 2 && same(i, i + 1);
    if (repeat) {
        n = 2;
        while (n < cap && same(i + n - 1, i + n)) ++n;
        return n;
    }
    while (n < cap && !(n <= 126 && i + n + 1 < npix && same(i + n, i + n + 1))) ++n;
    return n;
}

// WriteTgaFile + unloadRleData (src/tgaimage.cpp:172-300): bottom-left origin,
// greedy raw/run chunks of at most 128 pixels, 26-byte footer block.
bool tga_write(const std::string &path, int w, int h, int bpp, const uint8_t *data, bool rle)
{
    std::ofstream out(path, std::ios::binary);
    if (!out.is_open()) return false;
    TgaHeader hd;
    std::memset(&hd, 0, sizeof(hd));
    hd.bitsperpixel = (uint8_t)(bpp << 3);
    hd.width = (uint16_t)w;
    hd.height = (uint16_t)h;
    hd.datatypecode = (bpp == 1) ? (rle ? 11 : 3) : (rle ? 10 : 2);
    hd.imagedescriptor = 0x00;
    out.write(reinterpret_cast<const char *>(&hd), sizeof(hd));
    if (!rle) {
        out.write(reinterpret_cast<const char *>(data), (size_t)w * h * bpp);
    } else {
        const size_t npix = (size_t)w * h;
        for (size_t cur = 0; cur < npix;) {
            bool repeat = false;
            const size_t n = rle_packet(data, npix, bpp, cur, repeat);
            out.put((char)(repeat ? n + 127 : n - 1));
            out.write(reinterpret_cast<const char *>(data + cur * bpp), (repeat ? 1 : n) * (size_t)bpp);
            cur += n;
        }
    }
    const uint8_t zeros[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const char footer[18] = {'T', 'R', 'U', 'E', 'V', 'I', 'S', 'I', 'O',
                             'N', '-', 'X', 'F', 'I', 'L', 'E', '.', '\0'};
    out.write(reinterpret_cast<const char *>(zeros), 8);
    out.write(footer, 18);
    return out.good();
}

// ---------------- OBJ / MTL (src/model.cpp:16-476) ----------------

static std::string ltrim(const std::string &s)
{
    size_t p = s.find_first_not_of(" \n\r\t\f\v");
    return p == std::string::npos ? "" : s.substr(p);
}

static bool starts(const std::string &l, const char *k) { return l.compare(0, std::strlen(k), k) == 0; }

// istream >> float: whitespace-separated token, correctly rounded.
static bool read_float(std::istringstream &iss, float &v)
{
    std::string tok;
    if (!(iss >> tok)) return false;
    char *end = nullptr;
    float f = std::strtof(tok.c_str(), &end);
    if (end == tok.c_str()) return false;
    v = f;
    return true;
}

struct MtlEntry {
    fgl_material m;
};

struct LoadState {
    int wrap = 0, filter = 0;
    std::map<std::string, int> tex_cache;  // path|wrap|filter -> texture index
};

static int load_texture(HostScene &s, LoadState &st, const std::string &path)
{
    std::string key = path + "|" + std::to_string(st.wrap) + "|" + std::to_string(st.filter);
    auto it = st.tex_cache.find(key);
    if (it != st.tex_cache.end()) return it->second;
    HostTexture t;
    std::string err;
    if (!tga_read(path, t, err)) {
        std::fprintf(stderr, "[warning] Failed to load texture: %s (%s)\n", path.c_str(), err.c_str());
        st.tex_cache[key] = -1;
        return -1;
    }
    flip_vertically(t);  // flipTexCoordY (src/model.cpp:423)
    t.wrap = st.wrap;
    t.filter = st.filter;
    s.textures.push_back(std::move(t));
    int id = (int)s.textures.size() - 1;
    st.tex_cache[key] = id;
    return id;
}

static void default_material(fgl_material &m)
{
    std::memset(&m, 0, sizeof(m));
    m.map_kd = m.map_ks = m.map_ke = m.map_bump = -1;
    m.albedo[0] = m.albedo[1] = m.albedo[2] = 1.f;
    m.map_base = m.map_rough = m.map_metal = m.map_ao = m.map_pbr_ke = -1;
}

static void load_mtl(HostScene &s, LoadState &st, const std::string &dir, const std::string &file,
                     std::map<std::string, fgl_material> &mats)
{
    std::ifstream in(dir + file);
    if (in.fail()) {
        std::fprintf(stderr, "[error] Cannot open the .mtl file: '%s'\n", (dir + file).c_str());
        return;
    }
    std::string line, name;
    while (!in.eof()) {
        std::getline(in, line);
        line = ltrim(line);
        std::istringstream iss(line);
        std::string trash;
        auto cur = [&]() -> fgl_material * {
            auto it = mats.find(name);
            return it == mats.end() ? nullptr : &it->second;
        };
        auto vec = [&](float *dst) {
            iss >> trash;
            float v[3] = {0.f, 0.f, 0.f};
            for (int i = 0; i < 3; ++i)
                if (!read_float(iss, v[i])) break;
            dst[0] = v[0]; dst[1] = v[1]; dst[2] = v[2];
        };
        auto tex = [&]() -> int {
            std::string f;
            iss >> trash >> f;
            return load_texture(s, st, dir + f);
        };
        if (starts(line, "newmtl ")) {
            iss >> trash >> name;
            fgl_material m;
            default_material(m);
            mats[name] = m;
        } else if (!cur()) {
            continue;
        } else if (starts(line, "Ka ")) {
            vec(cur()->ka);
        } else if (starts(line, "Kd ")) {
            vec(cur()->kd);
            std::memcpy(cur()->albedo, cur()->kd, sizeof(float) * 3);
        } else if (starts(line, "Ks ")) {
            vec(cur()->ks);
        } else if (starts(line, "Ke ")) {
            vec(cur()->ke);
            std::memcpy(cur()->pbr_ke, cur()->ke, sizeof(float) * 3);
        } else if (starts(line, "Pr ")) {
            iss >> trash;
            float f = 0.f;
            read_float(iss, f);
            cur()->roughness = f;
        } else if (starts(line, "Pm ")) {
            iss >> trash;
            float f = 0.f;
            read_float(iss, f);
            cur()->metalness = f;
        } else if (starts(line, "map_Kd ")) {
            int id = tex();
            if (id >= 0) cur()->map_kd = cur()->map_base = id;
        } else if (starts(line, "map_Ks ")) {
            int id = tex();
            if (id >= 0) cur()->map_ks = id;
        } else if (starts(line, "map_Ke ")) {
            int id = tex();
            if (id >= 0) cur()->map_ke = cur()->map_pbr_ke = id;
        } else if (starts(line, "map_Bump ") || starts(line, "norm ")) {
            int id = tex();
            if (id >= 0) cur()->map_bump = id;
        } else if (starts(line, "map_Ao ")) {
            int id = tex();
            if (id >= 0) cur()->map_ao = id;
        } else if (starts(line, "map_Pr ")) {
            int id = tex();
            if (id >= 0) cur()->map_rough = id;
        } else if (starts(line, "map_Pm ")) {
            int id = tex();
            if (id >= 0) cur()->map_metal = id;
        }
    }
}

struct MeshBuild {
    std::vector<int32_t> vi, ti, ni;
    std::string material;
    bool has_material = false;
};

// Tangents (src/model.cpp:430-476): per face in mesh-map then face order,
// accumulated per position index, then normalised ((1,0,0) if zero).
static void generate_tangents(HostModel &m)
{
    size_t nv = m.pos.size() / 3;
    std::vector<Vec3> acc(nv, vec3(0.f, 0.f, 0.f));
    auto P = [&](int i) { return vec3(m.pos[3 * i], m.pos[3 * i + 1], m.pos[3 * i + 2]); };
    auto UV = [&](int i) { return vec2(m.uv[2 * i], m.uv[2 * i + 1]); };
    for (int f = 0; f < m.n_faces(); ++f) {
        int p0 = m.vi[3 * f], p1 = m.vi[3 * f + 1], p2 = m.vi[3 * f + 2];
        Vec3 e1 = sub3(P(p1), P(p0)), e2 = sub3(P(p2), P(p0));
        Vec2 t0 = UV(m.ti[3 * f]), t1 = UV(m.ti[3 * f + 1]), t2 = UV(m.ti[3 * f + 2]);
        Vec2 d1 = vec2(t1.x - t0.x, t1.y - t0.y), d2 = vec2(t2.x - t0.x, t2.y - t0.y);
        float det = d1.x * d2.y - d2.x * d1.y;
        if (det == 0.f) continue;
        float inv = 1.f / det;
        Vec3 T = norm3(scale3(vec3(d2.y * e1.x - d1.y * e2.x, d2.y * e1.y - d1.y * e2.y,
                                   d2.y * e1.z - d1.y * e2.z),
                              inv));
        acc[p0] = add3(acc[p0], T);
        acc[p1] = add3(acc[p1], T);
        acc[p2] = add3(acc[p2], T);
    }
    m.tan.resize(nv * 3);
    for (size_t i = 0; i < nv; ++i) {
        Vec3 v = acc[i];
        if (len3(v) == 0.f) v = vec3(1.f, 0.f, 0.f);
        else v = norm3(v);
        m.tan[3 * i] = v.x; m.tan[3 * i + 1] = v.y; m.tan[3 * i + 2] = v.z;
    }
}

// normalizePositionVertices (src/model.cpp:107-138).  Its initial maxima are
// FLT_MIN (smallest positive normal), and the translation is not scaled.
static void normalize_positions(HostModel &m)
{
    float xmin = FLT_MAX, xmax = FLT_MIN, ymin = FLT_MAX, ymax = FLT_MIN, zmin = FLT_MAX, zmax = FLT_MIN;
    size_t nv = m.pos.size() / 3;
    for (size_t i = 0; i < nv; ++i) {
        float x = m.pos[3 * i], y = m.pos[3 * i + 1], z = m.pos[3 * i + 2];
        xmin = smin(xmin, x); xmax = smax(xmax, x);
        ymin = smin(ymin, y); ymax = smax(ymax, y);
        zmin = smin(zmin, z); zmax = smax(zmax, z);
    }
    float sc = 2.f / smax(xmax - xmin, smax(ymax - ymin, zmax - zmin));
    float M[16];
    identity(M);
    M[0] = M[5] = M[10] = sc;
    M[3] = -(xmax + xmin) / (xmax - xmin);
    M[7] = -(ymax + ymin) / (ymax - ymin);
    M[11] = -(zmax + zmin) / (zmax - zmin);
    for (size_t i = 0; i < nv; ++i) {
        Vec4 r = mv4(M, vec4(m.pos[3 * i], m.pos[3 * i + 1], m.pos[3 * i + 2], 1.f));
        m.pos[3 * i] = r.x; m.pos[3 * i + 1] = r.y; m.pos[3 * i + 2] = r.z;
    }
}

static bool load_obj(HostScene &s, LoadState &st, const std::string &path, bool normalized,
                     bool gen_tangent, HostModel &m, std::string &err)
{
    std::ifstream in(path);
    if (in.fail()) { err = "Failed to open the .obj file: " + path; return false; }
    std::map<std::string, MeshBuild> meshes;
    std::map<std::string, fgl_material> mats;
    std::string line, mesh_name;
    bool have_mesh = false;
    size_t slash = path.find_last_of('/');
    std::string dir = slash == std::string::npos ? "" : path.substr(0, slash + 1);
    while (!in.eof()) {
        std::getline(in, line);
        line = ltrim(line);
        std::istringstream iss(line);
        std::string trash;
        if (starts(line, "mtllib ")) {
            std::string f;
            iss >> trash >> f;
            load_mtl(s, st, dir, f, mats);
        } else if (starts(line, "v ")) {
            iss >> trash;
            float v[3] = {0.f, 0.f, 0.f};
            for (int i = 0; i < 3; ++i)
                if (!read_float(iss, v[i])) break;
            m.pos.insert(m.pos.end(), v, v + 3);
        } else if (starts(line, "vt ")) {
            iss >> trash;
            float v[2] = {0.f, 0.f};
            for (int i = 0; i < 2; ++i)
                if (!read_float(iss, v[i])) break;
            m.uv.insert(m.uv.end(), v, v + 2);
        } else if (starts(line, "vn ")) {
            iss >> trash;
            float v[3] = {0.f, 0.f, 0.f};
            for (int i = 0; i < 3; ++i)
                if (!read_float(iss, v[i])) break;
            m.nrm.insert(m.nrm.end(), v, v + 3);
        } else if (starts(line, "g ")) {
            iss >> trash >> mesh_name;
            meshes[mesh_name] = MeshBuild();  // a repeated name starts the mesh over
            have_mesh = true;
        } else if (starts(line, "usemtl ")) {
            std::string mn;
            iss >> trash >> mn;
            if (!have_mesh) { err = "usemtl before any g line: " + path; return false; }
            meshes[mesh_name].material = mn;
            meshes[mesh_name].has_material = true;
        } else if (starts(line, "f ")) {
            if (!have_mesh) { err = "face before any g line: " + path; return false; }
            iss >> trash;
            std::vector<std::array<int, 3>> verts;
            std::string tok;
            while (iss >> tok) {
                unsigned v, t, n;
                char c1, c2;
                std::istringstream ts(tok);
                if (!(ts >> v >> c1 >> t >> c2 >> n)) break;
                verts.push_back({(int)(v - 1), (int)(t - 1), (int)(n - 1)});
            }
            MeshBuild &mb = meshes[mesh_name];
            for (size_t i = 1; i + 1 < verts.size(); ++i) {
                const std::array<int, 3> *fv[3] = {&verts[0], &verts[i], &verts[i + 1]};
                for (int k = 0; k < 3; ++k) {
                    mb.vi.push_back((*fv[k])[0]);
                    mb.ti.push_back((*fv[k])[1]);
                    mb.ni.push_back((*fv[k])[2]);
                }
            }
        }
    }
    // material table of this model (scene-global ids)
    std::map<std::string, int> mat_id;
    for (auto &kv : mats) {
        s.materials.push_back(kv.second);
        mat_id[kv.first] = (int)s.materials.size() - 1;
    }
    m.support_pbr = false;
    for (auto &kv : meshes) {
        MeshBuild &mb = kv.second;
        auto it = mb.has_material ? mat_id.find(mb.material) : mat_id.end();
        if (it == mat_id.end()) {
            err = "mesh '" + kv.first + "' has no material: " + path;
            return false;
        }
        int mid = it->second;
        for (size_t i = 0; i < mb.vi.size(); ++i) {
            m.vi.push_back(mb.vi[i]);
            m.ti.push_back(mb.ti[i]);
            m.ni.push_back(mb.ni[i]);
        }
        for (size_t f = 0; f < mb.vi.size() / 3; ++f) m.face_mat.push_back(mid);
        // SupportPBR is decided by the last mesh in map order (src/model.cpp:52-68)
        const fgl_material &fm = s.materials[mid];
        m.support_pbr = fm.map_metal >= 0 || fm.map_rough >= 0;
    }
    // bounds of every index stream
    int np = (int)m.pos.size() / 3, nt = (int)m.uv.size() / 2, nn = (int)m.nrm.size() / 3;
    for (size_t i = 0; i < m.vi.size(); ++i)
        if (m.vi[i] < 0 || m.vi[i] >= np || m.ti[i] < 0 || m.ti[i] >= nt || m.ni[i] < 0 || m.ni[i] >= nn) {
            err = "face index out of range: " + path;
            return false;
        }
    m.has_tangents = gen_tangent;
    if (gen_tangent) generate_tangents(m);
    if (normalized) normalize_positions(m);
    return true;
}

// ---------------- deterministic sample sets for the extensions ----------------

static uint64_t splitmix64(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static double urand(uint64_t &s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

// Vogel (golden-angle) disk of n points inside the unit disk.
static void vogel_disk(int n, float *out)
{
    const double golden = 2.39996322972865332;
    for (int i = 0; i < n; ++i) {
        double r = std::sqrt((i + 0.5) / n), th = i * golden;
        out[2 * i] = (float)(r * std::cos(th));
        out[2 * i + 1] = (float)(r * std::sin(th));
    }
}

static void ssao_kernel(fgl_ssao_cfg &c, int n, uint64_t seed)
{
    uint64_t s = seed;
    c.n_kernel = n;
    for (int i = 0; i < n; ++i) {
        double x = urand(s) * 2 - 1, y = urand(s) * 2 - 1, z = urand(s);
        double l = std::sqrt(x * x + y * y + z * z);
        if (l == 0) { x = 0; y = 0; z = 1; l = 1; }
        double u = urand(s);
        double sc = (double)i / n;
        sc = 0.1 + (1.0 - 0.1) * sc * sc;
        c.kernel[3 * i] = (float)(x / l * u * sc);
        c.kernel[3 * i + 1] = (float)(y / l * u * sc);
        c.kernel[3 * i + 2] = (float)(z / l * u * sc);
    }
    for (int i = 0; i < 16; ++i) {
        c.noise[3 * i] = (float)(urand(s) * 2 - 1);
        c.noise[3 * i + 1] = (float)(urand(s) * 2 - 1);
        c.noise[3 * i + 2] = 0.f;
    }
}

static void gaussian_weights(fgl_ssao_cfg &c, double sigma)
{
    int R = (int)std::ceil(3 * sigma);
    if (R > 8) R = 8;
    if (R < 1) R = 1;
    double w[9], sum = 0;
    for (int i = 0; i <= R; ++i) {
        w[i] = std::exp(-(double)i * i / (2 * sigma * sigma));
        sum += (i == 0) ? w[i] : 2 * w[i];
    }
    c.blur_radius = R;
    for (int i = 0; i < 9; ++i) c.blur_weights[i] = i <= R ? (float)(w[i] / sum) : 0.f;
}

static void set_shadow_pcss(fgl_shadow_cfg &c, int nb, int np, float area)
{
    c.mode = 2;
    c.n_blocker = nb;
    c.n_pcf = np;
    vogel_disk(nb, c.blocker_disk);
    vogel_disk(np, c.pcf_disk);
    c.blocker_filter = 0.005;  // PCSS_BLOCKER_SEARCH_FILTER_SIZE
    c.pcss_filter = 0.007;     // PCF_FILTER_SIZE
    c.pcf_filter = (float)0.007;
    c.area_light = area;
}

// ---------------- .scene (src/scene.cpp:19-196) ----------------

static int parse_wrap(const std::string &s)
{
    if (s == "repeat") return FGL_WRAP_REPEAT;
    if (s == "mirror") return FGL_WRAP_MIRROR;
    if (s == "clamp") return FGL_WRAP_CLAMP;
    return FGL_WRAP_NONE;
}

bool load_scene(const std::string &path, HostScene &s, std::string &err, int wrap, int filter)
{
    std::ifstream in(path);
    if (in.fail()) { err = "Failed to open the .scene file: " + path; return false; }
    fgl_scene_info &I = s.info;
    std::memset(&I, 0, sizeof(I));
    I.width = 1280;
    I.height = 800;
    I.ssaa_on = 0;
    I.ssaa_k = 2;
    I.ssao_on = 0;
    I.shadow_on = 1;     // Shadow's static default (src/shaders/shadow.cpp:9)
    I.deferred = 0;      // ForkerGL's static default render mode
    I.camera_ortho = 0;
    // The reference's own filters by default: SOFT_SHADOW_PCSS
    // (src/shaders/shadow.h:15-23) and Render::DoSSAO with its recursive blur,
    // drawing from the random stream as the reference does; the
    // `shadowfilter` and `ssaokernel` keys select the deterministic ones (the
    // kernel below, 32 taps of radius 0.075, is what `ssaokernel` replaces).
    set_shadow_pcss(I.shadow, 32, 64, 2.5f);
    I.shadow.mode = FGL_SHADOW_REF_PCSS;
    ssao_kernel(I.ssao, 32, 0x55A0u);
    I.ssao.radius = 0.075f;
    I.ssao.bias = 0.0005f;
    I.ssao.range = 0.01f;
    I.ssao.blur = 1;
    I.ssao.stochastic = 1;
    bool have_point = false, have_dir = false;
    fgl_light point{}, dir{};
    LoadState st;
    st.wrap = wrap;       // ForkerGL::TextureWrapping / TextureFiltering when the scene is
    st.filter = filter;   // loaded (src/forkergl.cpp:15-16, captured by Model::Load)
    std::map<std::string, std::shared_ptr<HostModel>> cache;
    std::string line;
    while (!in.eof()) {
        std::getline(in, line);
        line = ltrim(line);
        std::istringstream iss(line);
        std::string trash;
        if (starts(line, "#")) continue;
        if (starts(line, "mode ")) {
            std::string m;
            iss >> trash >> m;
            I.deferred = (m == "deferred");
        } else if (starts(line, "screen ")) {
            iss >> trash >> I.width >> I.height;
        } else if (starts(line, "ssaa ")) {
            std::string on;
            iss >> trash >> on >> I.ssaa_k;
            I.ssaa_on = (on == "on");
        } else if (starts(line, "ssao ")) {
            std::string on;
            iss >> trash >> on;
            I.ssao_on = (on == "on");
        } else if (starts(line, "shadow ")) {
            std::string on;
            iss >> trash >> on;
            I.shadow_on = (on == "on");
        } else if (starts(line, "light ")) {
            std::string type;
            iss >> trash >> type;
            fgl_light l{};
            for (int i = 0; i < 3; ++i) read_float(iss, l.vec[i]);
            for (int i = 0; i < 3; ++i) read_float(iss, l.color[i]);
            if (type == "point") { l.type = 0; point = l; have_point = true; }
            else if (type == "dir") { l.type = 1; dir = l; have_dir = true; }
        } else if (starts(line, "camera ")) {
            std::string type;
            iss >> trash >> type;
            if (type != "persp" && type != "ortho") continue;
            I.camera_ortho = (type == "ortho");
            for (int i = 0; i < 3; ++i) read_float(iss, I.eye[i]);
            for (int i = 0; i < 3; ++i) read_float(iss, I.look_at[i]);
        } else if (starts(line, "model ")) {
            std::string file, b1, b2;
            iss >> trash >> file >> b1 >> b2;
            bool normalized = b1 == "true", gen = b2 == "true";
            float p[3] = {0.f, 0.f, 0.f}, rot = 0.f, scale = 0.f;
            for (int i = 0; i < 3; ++i) read_float(iss, p[i]);
            read_float(iss, rot);
            read_float(iss, scale);
            // Identical model lines share one parsed model (the reference
            // re-parses the file; the arrays are the same).
            std::string key = file + "|" + b1 + "|" + b2 + "|" + std::to_string(st.wrap) + "|" +
                              std::to_string(st.filter);
            std::shared_ptr<HostModel> hm;
            auto it = cache.find(key);
            if (it != cache.end()) {
                hm = it->second;
            } else {
                hm = std::make_shared<HostModel>();
                if (!load_obj(s, st, file, normalized, gen, *hm, err)) return false;
                cache[key] = hm;
            }
            s.models.push_back(hm);
            std::array<float, 16> M;
            make_model_matrix(p[0], p[1], p[2], rot, scale, M.data());
            s.model_mats.push_back(M);
        }
        // ---- extension keys (ignored by the reference's parser) ----
        else if (starts(line, "shadowmap ")) {
            iss >> trash >> I.shadow_map_size;
        } else if (starts(line, "shadowfilter ")) {
            std::string kind;
            iss >> trash >> kind;
            if (kind == "hard") {
                I.shadow.mode = 0;
            } else if (kind == "grid") {
                I.shadow.mode = 3;
                iss >> I.shadow.grid;
            } else if (kind == "pcf") {
                int n = 64;
                float f = 0.007f;
                iss >> n;
                read_float(iss, f);
                I.shadow.mode = 1;
                I.shadow.n_pcf = n < 1 ? 1 : (n > FGL_MAX_DISK ? FGL_MAX_DISK : n);
                vogel_disk(I.shadow.n_pcf, I.shadow.pcf_disk);
                I.shadow.pcf_filter = f;
            } else if (kind == "pcss") {
                int nb = 32, np = 64;
                float area = 2.5f;
                iss >> nb >> np;
                read_float(iss, area);
                nb = nb < 1 ? 1 : (nb > FGL_MAX_DISK ? FGL_MAX_DISK : nb);
                np = np < 1 ? 1 : (np > FGL_MAX_DISK ? FGL_MAX_DISK : np);
                set_shadow_pcss(I.shadow, nb, np, area);
            } else if (kind == "refpcf" || kind == "refpcss") {
                // the reference's SOFT_SHADOW_PCF / SOFT_SHADOW_PCSS with its
                // random disk samples (src/shaders/shadow.h:14-23)
                set_shadow_pcss(I.shadow, 32, 64, 2.5f);
                I.shadow.mode = kind == "refpcf" ? FGL_SHADOW_REF_PCF : FGL_SHADOW_REF_PCSS;
            }
        } else if (starts(line, "ssaokernel reference")) {
            // Render::DoSSAO itself: RNG hemisphere samples from the stream
            I.ssao.stochastic = 1;
        } else if (starts(line, "ssaokernel ")) {
            int n = 64;
            float r = 0.5f, b = 0.0005f, rg = 0.01f;
            unsigned long long seed = 0x55A0u;
            I.ssao.stochastic = 0;
            iss >> trash >> n;
            read_float(iss, r);
            read_float(iss, b);
            read_float(iss, rg);
            iss >> seed;
            n = n < 1 ? 1 : (n > 64 ? 64 : n);
            ssao_kernel(I.ssao, n, seed);
            I.ssao.radius = r;
            I.ssao.bias = b;
            I.ssao.range = rg;
        } else if (starts(line, "aoblur ")) {
            std::string kind;
            iss >> trash >> kind;
            if (kind == "none") I.ssao.blur = 0;
            else if (kind == "reference") I.ssao.blur = 1;
            else if (kind == "gaussian") {
                float sigma = 2.f;
                read_float(iss, sigma);
                I.ssao.blur = 2;
                gaussian_weights(I.ssao, sigma);
            }
        } else if (starts(line, "texture ")) {
            std::string w, f;
            iss >> trash >> w >> f;
            st.wrap = parse_wrap(w);
            st.filter = (f == "linear") ? FGL_FILTER_LINEAR : FGL_FILTER_NEAREST;
        } else if (starts(line, "multilight ")) {
            std::string on;
            iss >> trash >> on;
            s.multilight = (on == "on");
        }
    }
    if (!have_point) {
        err = "scene has no point light: " + path;
        return false;
    }
    I.n_lights = 0;
    I.lights[I.n_lights++] = point;
    if (s.multilight && have_dir) I.lights[I.n_lights++] = dir;
    I.n_models = (int)s.models.size();
    I.n_materials = (int)s.materials.size();
    I.n_textures = (int)s.textures.size();
    return true;
}

}  // namespace fgl
