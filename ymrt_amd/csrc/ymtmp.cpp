// .ymtmp scene files -> ymrt_scene_desc.
//
// Follows SceneParser::parsing_scene (src/core/YMParser.cpp:2330-2433): lines
// lose every character outside the statement alphabet, '#' lines are dropped,
// the rest is joined and cut at ';'; every statement must be Begin(Type),
// End(Type) or attribute=value with a value of one of the accepted shapes
// (YMParser.cpp:420-762), blocks do not nest and hold only their type's
// attributes (765-814).  Blocks are then read in file order with the block
// readers' defaults and clamps (971-2231) into the containers SceneModule
// keeps (src/core/SceneModule.cpp): one name space for every named asset,
// references only to names declared earlier.  A block that fails a reader's
// check is skipped with a message and loading goes on, as in the reference.
//
// Transforms compose as InitTransformMatrix (src/Transform.h:36-60):
// parent * T * Rz * Ry * Rx * S with glm's translate / rotate / scale, matrix
// product and 4x4 inverse (external/glm 0.9.8 gtc/matrix_transform.inl,
// detail/type_mat4x4.inl, detail/func_matrix.inl) evaluated in the same
// operation order, so the matrices are the reference's floats.
#include "internal.hpp"

#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

namespace {

// ------------------------------------------------------------ glm-order math
struct M4 {
    float c[4][4];  // c[column][row], like glm::mat4
};

M4 m4_identity()
{
    M4 m{};
    for (int i = 0; i < 4; ++i) m.c[i][i] = 1.0f;
    return m;
}

// glm operator*(mat4, mat4): column i = A0*B[i][0] + A1*B[i][1] + A2*B[i][2] + A3*B[i][3]
M4 m4_mul(const M4 &a, const M4 &b)
{
    M4 r;
    for (int i = 0; i < 4; ++i)
        for (int k = 0; k < 4; ++k)
            r.c[i][k] = a.c[0][k] * b.c[i][0] + a.c[1][k] * b.c[i][1] + a.c[2][k] * b.c[i][2] + a.c[3][k] * b.c[i][3];
    return r;
}

// glm::translate(identity, v): column 3 = m0*v0 + m1*v1 + m2*v2 + m3
M4 m4_translate(const float v[3])
{
    const M4 m = m4_identity();
    M4 r = m;
    for (int k = 0; k < 4; ++k) r.c[3][k] = m.c[0][k] * v[0] + m.c[1][k] * v[1] + m.c[2][k] * v[2] + m.c[3][k];
    return r;
}

// glm::scale(identity, v): columns 0..2 scaled by v
M4 m4_scale(const float v[3])
{
    const M4 m = m4_identity();
    M4 r;
    for (int k = 0; k < 4; ++k) {
        r.c[0][k] = m.c[0][k] * v[0];
        r.c[1][k] = m.c[1][k] * v[1];
        r.c[2][k] = m.c[2][k] * v[2];
        r.c[3][k] = m.c[3][k];
    }
    return r;
}

// glm::rotate(identity, angle, axis) with glm::normalize(axis)
M4 m4_rotate(float angle, const float v[3])
{
    const float c = std::cos(angle);
    const float s = std::sin(angle);
    const float dot = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    const float inv = 1.0f / std::sqrt(dot);
    const float ax[3] = {v[0] * inv, v[1] * inv, v[2] * inv};
    const float t[3] = {(1.0f - c) * ax[0], (1.0f - c) * ax[1], (1.0f - c) * ax[2]};
    float R[3][3];
    R[0][0] = c + t[0] * ax[0];
    R[0][1] = t[0] * ax[1] + s * ax[2];
    R[0][2] = t[0] * ax[2] - s * ax[1];
    R[1][0] = t[1] * ax[0] - s * ax[2];
    R[1][1] = c + t[1] * ax[1];
    R[1][2] = t[1] * ax[2] + s * ax[0];
    R[2][0] = t[2] * ax[0] + s * ax[1];
    R[2][1] = t[2] * ax[1] - s * ax[0];
    R[2][2] = c + t[2] * ax[2];
    const M4 m = m4_identity();
    M4 r;
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 4; ++k) r.c[i][k] = m.c[0][k] * R[i][0] + m.c[1][k] * R[i][1] + m.c[2][k] * R[i][2];
    for (int k = 0; k < 4; ++k) r.c[3][k] = m.c[3][k];
    return r;
}

// glm compute_inverse<tmat4x4> (detail/func_matrix.inl:297-347)
M4 m4_inverse(const M4 &M)
{
    const auto &m = M.c;
    const float c00 = m[2][2] * m[3][3] - m[3][2] * m[2][3];
    const float c02 = m[1][2] * m[3][3] - m[3][2] * m[1][3];
    const float c03 = m[1][2] * m[2][3] - m[2][2] * m[1][3];
    const float c04 = m[2][1] * m[3][3] - m[3][1] * m[2][3];
    const float c06 = m[1][1] * m[3][3] - m[3][1] * m[1][3];
    const float c07 = m[1][1] * m[2][3] - m[2][1] * m[1][3];
    const float c08 = m[2][1] * m[3][2] - m[3][1] * m[2][2];
    const float c10 = m[1][1] * m[3][2] - m[3][1] * m[1][2];
    const float c11 = m[1][1] * m[2][2] - m[2][1] * m[1][2];
    const float c12 = m[2][0] * m[3][3] - m[3][0] * m[2][3];
    const float c14 = m[1][0] * m[3][3] - m[3][0] * m[1][3];
    const float c15 = m[1][0] * m[2][3] - m[2][0] * m[1][3];
    const float c16 = m[2][0] * m[3][2] - m[3][0] * m[2][2];
    const float c18 = m[1][0] * m[3][2] - m[3][0] * m[1][2];
    const float c19 = m[1][0] * m[2][2] - m[2][0] * m[1][2];
    const float c20 = m[2][0] * m[3][1] - m[3][0] * m[2][1];
    const float c22 = m[1][0] * m[3][1] - m[3][0] * m[1][1];
    const float c23 = m[1][0] * m[2][1] - m[2][0] * m[1][1];
    const float f0[4] = {c00, c00, c02, c03}, f1[4] = {c04, c04, c06, c07}, f2[4] = {c08, c08, c10, c11};
    const float f3[4] = {c12, c12, c14, c15}, f4[4] = {c16, c16, c18, c19}, f5[4] = {c20, c20, c22, c23};
    const float v0[4] = {m[1][0], m[0][0], m[0][0], m[0][0]};
    const float v1[4] = {m[1][1], m[0][1], m[0][1], m[0][1]};
    const float v2[4] = {m[1][2], m[0][2], m[0][2], m[0][2]};
    const float v3[4] = {m[1][3], m[0][3], m[0][3], m[0][3]};
    const float sa[4] = {+1.0f, -1.0f, +1.0f, -1.0f}, sb[4] = {-1.0f, +1.0f, -1.0f, +1.0f};
    M4 inv;
    for (int k = 0; k < 4; ++k) {
        inv.c[0][k] = (v1[k] * f0[k] - v2[k] * f1[k] + v3[k] * f2[k]) * sa[k];
        inv.c[1][k] = (v0[k] * f0[k] - v2[k] * f3[k] + v3[k] * f4[k]) * sb[k];
        inv.c[2][k] = (v0[k] * f1[k] - v1[k] * f3[k] + v3[k] * f5[k]) * sa[k];
        inv.c[3][k] = (v0[k] * f2[k] - v1[k] * f4[k] + v2[k] * f5[k]) * sb[k];
    }
    const float row0[4] = {inv.c[0][0], inv.c[1][0], inv.c[2][0], inv.c[3][0]};
    const float d[4] = {m[0][0] * row0[0], m[0][1] * row0[1], m[0][2] * row0[2], m[0][3] * row0[3]};
    const float det = (d[0] + d[1]) + (d[2] + d[3]);
    const float one_over = 1.0f / det;
    for (int i = 0; i < 4; ++i)
        for (int k = 0; k < 4; ++k) inv.c[i][k] = inv.c[i][k] * one_over;
    return inv;
}

float radians(float deg) { return deg * static_cast<float>(0.01745329251994329576923690768489); }

// ---------------------------------------------------------------- statements
bool keep_char(char c)
{
    return std::isalpha((unsigned char)c) || std::isdigit((unsigned char)c) || c == '=' || c == '<' || c == '>' ||
           c == ',' || c == '.' || c == '"' || c == '(' || c == ')' || c == '#' || c == '_' || c == ';' || c == '[' ||
           c == ']' || c == '+' || c == '-' || c == '\\' || c == '/' || c == ':';
}

const char *const kTypes[] = {"Camera",
                              "Transform",
                              "Mesh",
                              "Cube",
                              "Sphere",
                              "PrimitiveInstance",
                              "DefaultMaterial",
                              "LightMaterial",
                              "ImageTexture",
                              "ConstantTextureFloat",
                              "ConstantTextureRgb",
                              "CheckerBoardTexture",
                              "NoiseTexture",
                              "NoiseTextureFBM",
                              "NoiseTextureTurbulence",
                              "NoiseTextureMarble",
                              "NoiseTextureWood",
                              "NoiseTexturePolkaDot",
                              "NoiseTextureWave",
                              "DistantLight",
                              "Volume"};

// the attributes each block type accepts (is_*_attribute, YMParser.cpp:94-371)
const std::unordered_map<std::string, std::vector<std::string>> &attribute_sets()
{
    static const std::unordered_map<std::string, std::vector<std::string>> m = {
        {"Camera", {"from", "look", "fov"}},
        {"Transform", {"name", "translate", "scale", "rotate", "parent_transform_index"}},
        {"Mesh",
         {"name", "mesh_triangles", "mesh_position_indices", "mesh_positions", "mesh_normal_indices", "mesh_normals",
          "mesh_texcoord_indices", "mesh_texcoords"}},
        {"Cube", {"name"}},
        {"Sphere", {"name", "radius"}},
        {"PrimitiveInstance",
         {"name", "geometry_index", "transform_index", "material_index", "inner_volume_index", "treat_as_boundary"}},
        {"DefaultMaterial",
         {"name",
          "diffuse_albedo",
          "specular_albedo",
          "roughness_x",
          "roughness_y",
          "ior_n",
          "metalness",
          "specular_weight",
          "transmission_weight",
          "ior_priority",
          "coat_albedo",
          "coat_weight",
          "coat_thickness",
          "coat_ior",
          "coat_roughness_x",
          "coat_roughness_y",
          "diffuse_albedo_tex",
          "alpha_x_tex",
          "specular_albedo_tex",
          "alpha_y_tex",
          "specular_weight_tex",
          "metalness_tex",
          "transmission_weight_tex",
          "normal_mapping_tex",
          "bump_mapping_tex",
          "coat_albedo_tex",
          "coat_weight_tex",
          "coat_thickness_tex",
          "coat_roughness_x_tex",
          "coat_roughness_y_tex",
          "coat_normal_tex",
          "coat_normal_tex_type"}},
        {"LightMaterial", {"name", "light_color", "intensity"}},
        {"ImageTexture", {"name", "texture_file_name", "color_space"}},
        {"ConstantTextureFloat", {"name", "value"}},
        {"ConstantTextureRgb", {"name", "color"}},
        {"CheckerBoardTexture", {"name", "texture_black_idx", "texture_white_idx", "frequency"}},
        {"NoiseTexture", {"name", "texture_black_idx", "texture_white_idx", "frequency", "normalized"}},
        {"NoiseTextureFBM",
         {"name", "texture_black_idx", "texture_white_idx", "frequency", "lacunarity", "gain", "layer_count",
          "amplitude", "offset"}},
        {"NoiseTextureTurbulence",
         {"name", "texture_black_idx", "texture_white_idx", "frequency", "lacunarity", "gain", "layer_count",
          "amplitude", "offset"}},
        {"NoiseTextureMarble",
         {"name", "texture_black_idx", "texture_white_idx", "frequency", "lacunarity", "gain", "layer_count",
          "variation", "x_turb", "y_turb", "z_turb"}},
        {"NoiseTextureWood", {"name", "texture_black_idx", "texture_white_idx", "frequency", "variation"}},
        {"NoiseTexturePolkaDot", {"name", "texture_black_idx", "texture_white_idx", "frequency", "radius"}},
        {"NoiseTextureWave",
         {"name", "texture_black_idx", "texture_white_idx", "freq_0", "lacunarity_0", "gain_0", "layer_count_0",
          "freq_1", "lacunarity_1", "gain_1", "layer_count_1"}},
        {"DistantLight", {"name", "light_color", "transform_index", "intensity", "theta_max"}},
        {"Volume", {"name", "transform_index", "sigma_t", "albedo", "g", "density_texture_index"}},
    };
    return m;
}

bool is_type(const std::string &t)
{
    for (const char *k : kTypes)
        if (t == k) return true;
    return false;
}

bool digits(const std::string &s, size_t a, size_t b, int *count)
{
    for (size_t i = a; i < b; ++i)
        if (!std::isdigit((unsigned char)s[i])) return false;
    *count = (int)(b - a);
    return true;
}

// is_numeric_int (YMParser.cpp:420-437): digits only, no leading zero
bool is_int(const std::string &v)
{
    if (v.empty()) return false;
    if (v.size() > 1 && v[0] == '0') return false;
    int n;
    return digits(v, 0, v.size(), &n);
}

// is_numeric_float (439-529): [+-]digits, or [+-]digits?.digits?[f] with a digit somewhere
bool is_float(const std::string &v)
{
    if (v.empty()) return false;
    const size_t dot = v.find('.');
    auto signed_digits = [](const std::string &s, int *count) {
        *count = 0;
        for (size_t i = 0; i < s.size(); ++i) {
            const bool d = std::isdigit((unsigned char)s[i]) != 0;
            if (d) ++*count;
            if (i == 0 && !d && s[i] != '+' && s[i] != '-') return false;
            if (i > 0 && !d) return false;
        }
        return true;
    };
    if (dot == std::string::npos) {
        int n = 0;
        return signed_digits(v, &n) && n > 0;
    }
    const std::string l = v.substr(0, dot), r = v.substr(dot + 1);
    if (l.empty() && r.empty()) return false;
    int nl = 0, nr = 0;
    if (!l.empty() && !signed_digits(l, &nl)) return false;
    for (size_t i = 0; i < r.size(); ++i) {
        const bool d = std::isdigit((unsigned char)r[i]) != 0;
        if (d) ++nr;
        if (i + 1 == r.size() && !d && r[i] != 'f') return false;
        if (i + 1 < r.size() && !d) return false;
    }
    return nl + nr > 0;
}

// the comma-separated items of "<...>" or "[...]" (the item after the last
// separator ends at the closing bracket)
std::vector<std::string> items(const std::string &v, char close)
{
    std::vector<std::string> out;
    size_t start = 1;
    for (size_t end = 1; start < v.size() && end < v.size(); ++end) {
        if (v[end] != ',' && v[end] != close) continue;
        out.push_back(v.substr(start, end - start));
        start = end + 1;
    }
    return out;
}

bool is_vector(const std::string &v)
{
    if (v.size() < 2 || v.front() != '<' || v.back() != '>') return false;
    const auto it = items(v, '>');
    for (const auto &x : it)
        if (!is_float(x)) return false;
    return it.size() == 3;
}

bool quoted(const std::string &v, bool path)
{
    if (v.size() < 2 || v.front() != '"' || v.back() != '"') return false;
    for (size_t i = 1; i + 1 < v.size(); ++i) {
        const char c = v[i];
        const bool ok = std::isalpha((unsigned char)c) || std::isdigit((unsigned char)c) || c == '_' || c == '.' ||
                        c == '-' || c == '+' || (path && (c == '\\' || c == '/' || c == ':'));
        if (!ok) return false;
    }
    return true;
}

bool is_array(const std::string &v)
{
    if (v.size() < 2 || v.front() != '[' || v.back() != ']') return false;
    for (const auto &x : items(v, ']'))
        if (!(is_float(x) || is_int(x))) return false;
    return true;
}

bool is_value(const std::string &v)
{
    return is_float(v) || is_vector(v) || quoted(v, false) || quoted(v, true) || v == "true" || v == "false" ||
           is_int(v) || is_array(v);
}

struct Stmt {
    bool bracket;       // Begin/End(Type) or attribute=value
    std::string left;   // "Begin"/"End" or the attribute
    std::string right;  // the type or the value
    size_t line;        // statement number
};

// is_type_bracket / is_attribute_assign (630-734)
bool parse_statement(const std::string &s, size_t line, Stmt *out)
{
    const size_t lb = s.find('('), rb = s.rfind(')');
    if (lb != std::string::npos && rb != std::string::npos && lb < rb && rb == s.size() - 1) {
        const std::string type = s.substr(lb + 1, rb - lb - 1), be = s.substr(0, lb);
        if (is_type(type) && (be == "Begin" || be == "End")) {
            *out = Stmt{true, be, type, line};
            return true;
        }
    }
    const size_t eq = s.find('=');
    if (eq == std::string::npos) return false;
    const std::string attr = s.substr(0, eq), value = s.substr(eq + 1);
    for (char c : attr)
        if (!(std::isalpha((unsigned char)c) || c == '_')) return false;
    if (!is_value(value)) return false;
    *out = Stmt{false, attr, value, line};
    return true;
}

// ----------------------------------------------------------------- PNG input
// 8-bit PNG (grey, grey+alpha, RGB, RGBA, palette), not interlaced, decoded
// like stbi_load(..., 0) with vertical flip (LoadPng8, src/FileIO.cpp:17-45).
uint32_t be32(const uint8_t *p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

std::string decode_png(const std::string &path, uint32_t *w, uint32_t *h, uint32_t *ch, std::vector<uint8_t> *out)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) return "cannot open " + path;
    std::vector<uint8_t> d((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (d.size() < 8 || std::memcmp(d.data(), sig, 8) != 0) return path + ": not a PNG file";
    size_t p = 8;
    uint32_t width = 0, height = 0;
    int depth = 0, ctype = -1, interlace = 0;
    std::vector<uint8_t> idat, plte, trns;
    while (p + 8 <= d.size()) {
        const uint32_t len = be32(&d[p]);
        const std::string type(reinterpret_cast<const char *>(&d[p + 4]), 4);
        if (p + 12 + (size_t)len > d.size()) return path + ": truncated chunk";
        const uint8_t *c = &d[p + 8];
        if (type == "IHDR") {
            if (len < 13) return path + ": bad IHDR";
            width = be32(c);
            height = be32(c + 4);
            depth = c[8];
            ctype = c[9];
            interlace = c[12];
        } else if (type == "PLTE") {
            plte.assign(c, c + len);
        } else if (type == "tRNS") {
            trns.assign(c, c + len);
        } else if (type == "IDAT") {
            idat.insert(idat.end(), c, c + len);
        } else if (type == "IEND") {
            break;
        }
        p += 12 + (size_t)len;
    }
    if (width == 0 || height == 0) return path + ": empty image";
    if (depth != 8) return path + ": only 8-bit PNG is supported";
    if (interlace != 0) return path + ": interlaced PNG is not supported";
    int src_ch;
    switch (ctype) {
    case 0: src_ch = 1; break;
    case 2: src_ch = 3; break;
    case 3: src_ch = 1; break;
    case 4: src_ch = 2; break;
    case 6: src_ch = 4; break;
    default: return path + ": unknown PNG colour type";
    }
    if ((uint64_t)width * height > (1ull << 28)) return path + ": image too large";
    const size_t stride = (size_t)width * src_ch;
    std::vector<uint8_t> raw((stride + 1) * height);
    uLongf raw_len = (uLongf)raw.size();
    if (uncompress(raw.data(), &raw_len, idat.data(), (uLong)idat.size()) != Z_OK || raw_len != raw.size())
        return path + ": corrupt image data";
    std::vector<uint8_t> img(stride * height);
    for (uint32_t y = 0; y < height; ++y) {
        const uint8_t ft = raw[y * (stride + 1)];
        const uint8_t *src = &raw[y * (stride + 1) + 1];
        uint8_t *dst = &img[y * stride];
        const uint8_t *prev = y ? &img[(y - 1) * stride] : nullptr;
        for (size_t x = 0; x < stride; ++x) {
            const int a = x >= (size_t)src_ch ? dst[x - src_ch] : 0;
            const int b = prev ? prev[x] : 0;
            const int cc = (prev && x >= (size_t)src_ch) ? prev[x - src_ch] : 0;
            int v = src[x];
            switch (ft) {
            case 0: break;
            case 1: v += a; break;
            case 2: v += b; break;
            case 3: v += (a + b) / 2; break;
            case 4: {
                const int pp = a + b - cc, pa = std::abs(pp - a), pb = std::abs(pp - b), pc = std::abs(pp - cc);
                v += (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : cc);
                break;
            }
            default: return path + ": bad PNG filter";
            }
            dst[x] = (uint8_t)v;
        }
    }
    uint32_t out_ch = (uint32_t)src_ch;
    if (ctype == 3) out_ch = trns.empty() ? 3 : 4;
    out->assign((size_t)width * height * out_ch, 0);
    for (uint32_t y = 0; y < height; ++y) {
        const uint32_t sy = height - 1 - y;  // row 0 of the texture is the bottom row
        for (uint32_t x = 0; x < width; ++x) {
            uint8_t *o = &(*out)[((size_t)y * width + x) * out_ch];
            const uint8_t *s = &img[(size_t)sy * stride + (size_t)x * src_ch];
            if (ctype == 3) {
                const uint32_t k = s[0];
                if (3 * k + 2 >= plte.size()) return path + ": palette index out of range";
                o[0] = plte[3 * k];
                o[1] = plte[3 * k + 1];
                o[2] = plte[3 * k + 2];
                if (out_ch == 4) o[3] = k < trns.size() ? trns[k] : 255;
            } else {
                std::memcpy(o, s, out_ch);
            }
        }
    }
    *w = width;
    *h = height;
    *ch = out_ch;
    return "";
}

// --------------------------------------------------------------- built-ins
// The Cube geometry (GeometryData::InitCube, src/GeometryDefines.h:274-519):
// 12 triangles of 36 corners, corner k of triangle t at index 3t+k.  Each
// face is two triangles over its corner sequence (a b c, d e f).
struct CubeFace {
    signed char p[6][3];
    signed char n[3];
    signed char uv[6][2];
};
const CubeFace kCube[6] = {
    {{{-1, -1, -1}, {1, 1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, -1, -1}, {-1, 1, -1}},
     {0, 0, -1},
     {{0, 1}, {1, 0}, {1, 1}, {1, 0}, {0, 1}, {0, 0}}},
    {{{-1, -1, 1}, {1, -1, 1}, {1, 1, 1}, {1, 1, 1}, {-1, 1, 1}, {-1, -1, 1}},
     {0, 0, 1},
     {{0, 0}, {1, 0}, {1, 1}, {1, 1}, {0, 1}, {0, 0}}},
    {{{-1, 1, 1}, {-1, 1, -1}, {-1, -1, -1}, {-1, -1, -1}, {-1, -1, 1}, {-1, 1, 1}},
     {-1, 0, 0},
     {{1, 0}, {1, 1}, {0, 1}, {0, 1}, {0, 0}, {1, 0}}},
    {{{1, 1, 1}, {1, -1, -1}, {1, 1, -1}, {1, -1, -1}, {1, 1, 1}, {1, -1, 1}},
     {1, 0, 0},
     {{1, 1}, {0, 0}, {1, 0}, {0, 0}, {1, 1}, {0, 1}}},
    {{{-1, -1, -1}, {1, -1, -1}, {1, -1, 1}, {1, -1, 1}, {-1, -1, 1}, {-1, -1, -1}},
     {0, -1, 0},
     {{0, 0}, {1, 0}, {1, 1}, {1, 1}, {0, 1}, {0, 0}}},
    {{{-1, 1, -1}, {1, 1, 1}, {1, 1, -1}, {1, 1, 1}, {-1, 1, -1}, {-1, 1, 1}},
     {0, 1, 0},
     {{0, 1}, {1, 0}, {1, 1}, {1, 0}, {0, 1}, {0, 0}}},
};

}  // namespace

// ------------------------------------------------------------------- loader
struct ymrt_ymtmp {
    enum Kind { kTransform, kMesh, kSphere, kMaterial, kTexture, kInstance, kDistant, kVolume };
    struct Mesh {
        std::vector<uint32_t> pidx, nidx, tidx;
        std::vector<float> pos, nrm, uv;
    };
    struct Named {
        Kind kind;
        uint32_t index;  // YMRT_INVALID_INDEX when the asset failed to load
    };

    std::vector<Mesh> meshes;
    std::vector<ymrt_mesh_desc> mesh_descs;
    std::vector<ymrt_sphere_desc> spheres;
    std::vector<ymrt_instance_desc> instances;
    std::vector<ymrt_material_desc> materials;
    std::vector<ymrt_texture_desc> textures;
    std::vector<std::vector<uint8_t>> texels;
    std::vector<ymrt_distant_light_desc> distant;
    std::vector<ymrt_volume_desc> volumes;
    std::vector<M4> transforms, inverses;
    std::vector<std::string> warnings;
    std::vector<std::string> names_in_order;
    std::unordered_map<std::string, Named> names;
    ymrt_camera_desc camera{};
    std::string scene_path;
    ymrt_scene_desc desc{};

    void finish();
};

namespace {

using Y = ymrt_ymtmp;

const char *kind_name(Y::Kind k)
{
    switch (k) {
    case Y::kTransform: return "Transform";
    case Y::kMesh:
    case Y::kSphere: return "geometry";
    case Y::kMaterial: return "material";
    case Y::kTexture: return "texture";
    case Y::kInstance: return "PrimitiveInstance";
    case Y::kDistant: return "DistantLight";
    default: return "Volume";
    }
}

// One block's statements with typed accessors.  A failed accessor records a
// message for the block and returns false; the reader then skips the block.
struct Block {
    Y &y;
    const std::vector<Stmt> &st;
    size_t begin, end;  // statements begin+1 .. end-1
    std::string type, err;

    bool fail(const Stmt &s, const std::string &m)
    {
        err = m + " (" + type + " block, statement " + std::to_string(s.line) + ")";
        return false;
    }
    static bool to_float(const std::string &v, float *f)
    {
        try {
            *f = std::stof(v);
            return true;
        } catch (const std::exception &) {
            return false;
        }
    }
    bool get_float(const Stmt &s, float *f)
    {
        if (!is_float(s.right) && !is_int(s.right)) return fail(s, "expected a number for " + s.left);
        if (!to_float(s.right, f)) return fail(s, "number out of range for " + s.left);
        return true;
    }
    bool get_int(const Stmt &s, int *i)
    {
        if (!is_float(s.right) && !is_int(s.right)) return fail(s, "expected an integer for " + s.left);
        try {
            *i = std::stoi(s.right);
        } catch (const std::exception &) {
            return fail(s, "integer out of range for " + s.left);
        }
        return true;
    }
    bool get_vec(const Stmt &s, float v[3])
    {
        if (!is_vector(s.right)) return fail(s, "expected <x,y,z> for " + s.left);
        const auto it = items(s.right, '>');
        for (int k = 0; k < 3; ++k)
            if (!to_float(it[k], &v[k])) return fail(s, "number out of range for " + s.left);
        return true;
    }
    static bool get_bool(const Stmt &s) { return s.right == "true"; }
    bool get_string(const Stmt &s, std::string *out)
    {
        if (s.right.size() < 2 || s.right.front() != '"' || s.right.back() != '"')
            return fail(s, "expected a quoted string for " + s.left);
        *out = s.right.substr(1, s.right.size() - 2);
        return true;
    }
    // attribute_get_name (885-897): a new name
    bool get_name(const Stmt &s, std::string *out)
    {
        if (!get_string(s, out)) return false;
        if (y.names.count(*out)) return fail(s, "name \"" + *out + "\" is already declared");
        return true;
    }
    // attribute_get_index_from_name (899-913) restricted to the expected kinds
    bool get_ref(const Stmt &s, Y::Kind k0, Y::Kind k1, uint32_t *index, Y::Kind *kind = nullptr)
    {
        std::string n;
        if (!get_string(s, &n)) return false;
        const auto it = y.names.find(n);
        if (it == y.names.end()) return fail(s, "undeclared name \"" + n + "\"");
        if (it->second.kind != k0 && it->second.kind != k1)
            return fail(s, "\"" + n + "\" names a " + std::string(kind_name(it->second.kind)) + ", not a " +
                               kind_name(k0));
        *index = it->second.index;
        if (kind) *kind = it->second.kind;
        return true;
    }
    bool get_ref(const Stmt &s, Y::Kind k, uint32_t *index) { return get_ref(s, k, k, index); }
    bool get_int_array(const Stmt &s, std::vector<uint32_t> *out)
    {
        if (!is_array(s.right)) return fail(s, "expected [..] for " + s.left);
        for (const auto &x : items(s.right, ']')) {
            try {
                out->push_back((uint32_t)std::stoi(x));
            } catch (const std::exception &) {
                return fail(s, "integer out of range in " + s.left);
            }
        }
        return true;
    }
    bool get_float_array(const Stmt &s, std::vector<float> *out)
    {
        if (!is_array(s.right)) return fail(s, "expected [..] for " + s.left);
        for (const auto &x : items(s.right, ']')) {
            float f;
            if (!to_float(x, &f)) return fail(s, "number out of range in " + s.left);
            out->push_back(f);
        }
        return true;
    }
    bool named(const std::string &name)
    {
        if (name.empty()) {
            err = type + " block requires a name (statement " + std::to_string(st[begin].line) + ")";
            return false;
        }
        return true;
    }
};

void add_name(Y &y, const std::string &n, Y::Kind k, uint32_t index)
{
    y.names[n] = Y::Named{k, index};
    y.names_in_order.push_back(n);
}

float maxf(float a, float b) { return a < b ? b : a; }  // glm::max(a, b): a < b ? b : a
float clampf(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }

ymrt_texture_desc texture_base(uint32_t type)
{
    ymrt_texture_desc t{};
    t.type = type;
    t.child[0] = t.child[1] = YMRT_INVALID_INDEX;
    t.scale_u = t.scale_v = 1.0f;
    return t;
}

// -------------------------------------------------------------- block readers
bool read_camera(Block &b)
{
    float from[3] = {0.0f, 0.0f, 10.0f}, look[3] = {0.0f, 0.0f, 4.0f}, fov = 45.0f;
    for (size_t i = b.begin + 1; i < b.end; ++i) {
        const Stmt &s = b.st[i];
        if (s.left == "from") {
            if (!b.get_vec(s, from)) return false;
        } else if (s.left == "look") {
            if (!b.get_vec(s, look)) return false;
        } else if (s.left == "fov") {
            if (!b.get_float(s, &fov)) return false;
            fov = maxf(fov, 0.01f);
        }
    }
    ymrt_camera_desc &c = b.y.camera;
    for (int k = 0; k < 3; ++k) {
        c.position[k] = from[k];
        c.w[k] = from[k] - look[k];
    }
    c.fov = radians(fov);
    return true;
}

bool read_transform(Block &b)
{
    std::string name;
    float t[3] = {0.0f, 0.0f, 0.0f}, sc[3] = {1.0f, 1.0f, 1.0f}, r[3] = {0.0f, 0.0f, 0.0f};
    uint32_t parent = YMRT_INVALID_INDEX;
    for (size_t i = b.begin + 1; i < b.end; ++i) {
        const Stmt &s = b.st[i];
        if (s.left == "name") {
            if (!b.get_name(s, &name)) return false;
        } else if (s.left == "translate") {
            if (!b.get_vec(s, t)) return false;
        } else if (s.left == "scale") {
            if (!b.get_vec(s, sc)) return false;
            for (float &v : sc) v = maxf(0.0001f, v);
        } else if (s.left == "rotate") {
            if (!b.get_vec(s, r)) return false;
            for (float &v : r) v = clampf(v, 0.0f, 360.0f);
        } else if (s.left == "parent_transform_index") {
            if (!b.get_ref(s, Y::kTransform, &parent)) return false;
        }
    }
    if (!b.named(name)) return false;
    static const float ax[3] = {1.0f, 0.0f, 0.0f}, ay[3] = {0.0f, 1.0f, 0.0f}, az[3] = {0.0f, 0.0f, 1.0f};
    M4 m = m4_mul(m4_mul(m4_mul(m4_mul(m4_translate(t), m4_rotate(radians(r[2]), az)), m4_rotate(radians(r[1]), ay)),
                         m4_rotate(radians(r[0]), ax)),
                  m4_scale(sc));
    if (parent != YMRT_INVALID_INDEX) m = m4_mul(b.y.transforms[parent], m);
    b.y.transforms.push_back(m);
    b.y.inverses.push_back(m4_inverse(m));
    add_name(b.y, name, Y::kTransform, (uint32_t)b.y.transforms.size() - 1);
    return true;
}

bool read_mesh(Block &b)
{
    std::string name;
    std::vector<uint32_t> tris, pidx, nidx, tidx;
    std::vector<float> pos, nrm, uv;
    for (size_t i = b.begin + 1; i < b.end; ++i) {
        const Stmt &s = b.st[i];
        bool ok = true;
        if (s.left == "name") ok = b.get_name(s, &name);
        else if (s.left == "mesh_triangles") ok = b.get_int_array(s, &tris);
        else if (s.left == "mesh_position_indices") ok = b.get_int_array(s, &pidx);
        else if (s.left == "mesh_positions") ok = b.get_float_array(s, &pos);
        else if (s.left == "mesh_normal_indices") ok = b.get_int_array(s, &nidx);
        else if (s.left == "mesh_normals") ok = b.get_float_array(s, &nrm);
        else if (s.left == "mesh_texcoord_indices") ok = b.get_int_array(s, &tidx);
        else if (s.left == "mesh_texcoords") ok = b.get_float_array(s, &uv);
        if (!ok) return false;
    }
    if (!b.named(name)) return false;
    tris.resize(tris.size() / 3 * 3);  // attribute_get_triangle_array drops a partial triple
    const size_t nt = tris.size() / 3;
    auto err = [&](const char *m) {
        b.err = std::string("mesh \"") + name + "\": " + m;
        return false;
    };
    if (pidx.size() != nt * 3) return err("position index count is not 3 per triangle");
    if (nidx.size() != nt * 3) return err("normal index count is not 3 per triangle");
    if (!tidx.empty() && tidx.size() != nt * 3) return err("texcoord index count is not 3 per triangle");
    if (nrm.empty() || pos.empty()) return err("positions and normals are required");
    if (nt == 0) return err("no triangles");
    // corner k of triangle t reads the per-corner index arrays at tris[3t+k]
    // (FetchMeshShadingData, src/ShadingUtilities.cuh:733-756)
    Y::Mesh m;
    const size_t np = pos.size() / 3, nn = nrm.size() / 3, nu = uv.size() / 2;
    for (size_t c = 0; c < nt * 3; ++c) {
        const uint32_t k = tris[c];
        if (k >= pidx.size() || pidx[k] >= np || k >= nidx.size() || nidx[k] >= nn)
            return err("triangle corner index out of range");
        m.pidx.push_back(pidx[k]);
        m.nidx.push_back(nidx[k]);
        if (!tidx.empty()) {
            if (k >= tidx.size() || tidx[k] >= nu) return err("texcoord index out of range");
            m.tidx.push_back(tidx[k]);
        }
    }
    m.pos.assign(pos.begin(), pos.begin() + np * 3);
    m.nrm.assign(nrm.begin(), nrm.begin() + nn * 3);
    if (!m.tidx.empty()) m.uv.assign(uv.begin(), uv.begin() + nu * 2);
    b.y.meshes.push_back(std::move(m));
    add_name(b.y, name, Y::kMesh, (uint32_t)b.y.meshes.size() - 1);
    return true;
}

bool read_cube(Block &b)
{
    std::string name;
    for (size_t i = b.begin + 1; i < b.end; ++i)
        if (b.st[i].left == "name" && !b.get_name(b.st[i], &name)) return false;
    if (!b.named(name)) return false;
    Y::Mesh m;
    for (const CubeFace &f : kCube) {
        for (int k = 0; k < 6; ++k) {
            const uint32_t c = (uint32_t)m.pidx.size();
            for (int a = 0; a < 3; ++a) m.pos.push_back(0.5f * f.p[k][a]);
            for (int a = 0; a < 3; ++a) m.nrm.push_back((float)f.n[a]);
            for (int a = 0; a < 2; ++a) m.uv.push_back((float)f.uv[k][a]);
            m.pidx.push_back(c);
            m.nidx.push_back(c);
            m.tidx.push_back(c);
        }
    }
    b.y.meshes.push_back(std::move(m));
    add_name(b.y, name, Y::kMesh, (uint32_t)b.y.meshes.size() - 1);
    return true;
}

bool read_sphere(Block &b)
{
    std::string name;
    float radius = 1.0f;
    for (size_t i = b.begin + 1; i < b.end; ++i) {
        const Stmt &s = b.st[i];
        if (s.left == "name") {
            if (!b.get_name(s, &name)) return false;
        } else if (s.left == "radius") {
            if (!b.get_float(s, &radius)) return false;
            radius = maxf(0.0001f, radius);
        }
    }
    if (!b.named(name)) return false;
    b.y.spheres.push_back(ymrt_sphere_desc{radius});
    add_name(b.y, name, Y::kSphere, (uint32_t)b.y.spheres.size() - 1);
    return true;
}

bool read_instance(Block &b)
{
    std::string name;
    uint32_t geom = YMRT_INVALID_INDEX, xf = YMRT_INVALID_INDEX, mat = YMRT_INVALID_INDEX, vol = YMRT_INVALID_INDEX;
    Y::Kind gkind = Y::kMesh;
    bool boundary = false;
    for (size_t i = b.begin + 1; i < b.end; ++i) {
        const Stmt &s = b.st[i];
        bool ok = true;
        if (s.left == "name") ok = b.get_name(s, &name);
        else if (s.left == "geometry_index") ok = b.get_ref(s, Y::kMesh, Y::kSphere, &geom, &gkind);
        else if (s.left == "transform_index") ok = b.get_ref(s, Y::kTransform, &xf);
        else if (s.left == "material_index") ok = b.get_ref(s, Y::kMaterial, &mat);
        else if (s.left == "inner_volume_index") ok = b.get_ref(s, Y::kVolume, &vol);
        else if (s.left == "treat_as_boundary") boundary = Block::get_bool(s);
        if (!ok) return false;
    }
    if (!b.named(name)) return false;
    if (geom == YMRT_INVALID_INDEX) {
        b.err = "PrimitiveInstance \"" + name + "\" requires a geometry";
        return false;
    }
    if (xf == YMRT_INVALID_INDEX) {
        b.err = "PrimitiveInstance \"" + name + "\" requires a transform";
        return false;
    }
    ymrt_instance_desc d{};
    d.geometry = geom;
    d.geometry_type = gkind == Y::kSphere ? YMRT_GEOMETRY_SPHERE : YMRT_GEOMETRY_MESH;
    d.material = mat;
    d.inner_volume = vol == YMRT_INVALID_INDEX ? -1 : (int32_t)vol;
    d.treat_as_boundary = boundary ? 1 : 0;
    std::memcpy(d.object_to_world, b.y.transforms[xf].c, 64);
    std::memcpy(d.world_to_object, b.y.inverses[xf].c, 64);
    b.y.instances.push_back(d);
    add_name(b.y, name, Y::kInstance, (uint32_t)b.y.instances.size() - 1);
    return true;
}

bool read_default_material(Block &b)
{
    std::string name;
    ymrt_material_desc m{};
    m.type = YMRT_MATERIAL_DEFAULT;
    float da[3] = {0.5f, 0.5f, 0.5f}, sa[3] = {1.0f, 1.0f, 1.0f}, ca[3] = {1.0f, 1.0f, 1.0f};
    m.roughness_x = m.roughness_y = 0.2f;
    m.ior_n = 1.3f;
    m.specular_weight = 1.0f;
    m.coat_thickness = 1.0f;
    m.coat_ior = 1.6f;
    m.coat_roughness_x = m.coat_roughness_y = 0.2f;
    uint32_t *texs[] = {&m.diffuse_albedo_tex,    &m.specular_albedo_tex,  &m.roughness_x_tex,
                        &m.roughness_y_tex,       &m.specular_weight_tex,  &m.metalness_tex,
                        &m.transmission_weight_tex, &m.coat_albedo_tex,    &m.coat_weight_tex,
                        &m.coat_thickness_tex,    &m.coat_roughness_x_tex, &m.coat_roughness_y_tex,
                        &m.normal_mapping_tex,    &m.bump_mapping_tex};
    for (uint32_t *t : texs) *t = YMRT_INVALID_INDEX;
    const std::unordered_map<std::string, uint32_t *> tex_attr = {
        {"diffuse_albedo_tex", &m.diffuse_albedo_tex},
        {"alpha_x_tex", &m.roughness_x_tex},
        {"specular_albedo_tex", &m.specular_albedo_tex},
        {"alpha_y_tex", &m.roughness_y_tex},
        {"specular_weight_tex", &m.specular_weight_tex},
        {"metalness_tex", &m.metalness_tex},
        {"transmission_weight_tex", &m.transmission_weight_tex},
        {"normal_mapping_tex", &m.normal_mapping_tex},
        {"bump_mapping_tex", &m.bump_mapping_tex},
        {"coat_albedo_tex", &m.coat_albedo_tex},
        {"coat_weight_tex", &m.coat_weight_tex},
        {"coat_thickness_tex", &m.coat_thickness_tex},
        {"coat_roughness_x_tex", &m.coat_roughness_x_tex},
        {"coat_roughness_y_tex", &m.coat_roughness_y_tex},
    };
    uint32_t coat_normal_tex = YMRT_INVALID_INDEX;  // read and checked; shading does not use it
    for (size_t i = b.begin + 1; i < b.end; ++i) {
        const Stmt &s = b.st[i];
        const std::string &a = s.left;
        bool ok = true;
        float v;
        if (a == "name") ok = b.get_name(s, &name);
        else if (a == "diffuse_albedo") {
            ok = b.get_vec(s, da);
            for (float &x : da) x = maxf(x, 0.0f);
        } else if (a == "specular_albedo") {
            ok = b.get_vec(s, sa);
            for (float &x : sa) x = maxf(x, 0.0f);
        } else if (a == "coat_albedo") {
            ok = b.get_vec(s, ca);
            for (float &x : ca) x = maxf(x, 0.0f);
        } else if (a == "roughness_x" && (ok = b.get_float(s, &v))) m.roughness_x = clampf(v, 0.0f, 1.0f);
        else if (a == "roughness_y" && (ok = b.get_float(s, &v))) m.roughness_y = clampf(v, 0.0f, 1.0f);
        else if (a == "ior_n" && (ok = b.get_float(s, &v))) m.ior_n = maxf(0.0001f, v);
        else if (a == "metalness" && (ok = b.get_float(s, &v))) m.metalness = clampf(v, 0.0f, 1.0f);
        else if (a == "specular_weight" && (ok = b.get_float(s, &v))) m.specular_weight = clampf(v, 0.0f, 1.0f);
        else if (a == "transmission_weight" && (ok = b.get_float(s, &v)))
            m.transmission_weight = clampf(v, 0.0f, 1.0f);
        else if (a == "ior_priority") {
            int p = 0;
            if ((ok = b.get_int(s, &p))) m.ior_priority = (uint32_t)(p < 0 ? 0 : p);
        } else if (a == "coat_weight" && (ok = b.get_float(s, &v))) m.coat_weight = clampf(v, 0.0f, 1.0f);
        else if (a == "coat_thickness" && (ok = b.get_float(s, &v))) m.coat_thickness = maxf(v, 0.0f);
        else if (a == "coat_ior" && (ok = b.get_float(s, &v))) m.coat_ior = maxf(v, 0.0001f);
        else if (a == "coat_roughness_x" && (ok = b.get_float(s, &v))) m.coat_roughness_x = clampf(v, 0.0f, 1.0f);
        else if (a == "coat_roughness_y" && (ok = b.get_float(s, &v))) m.coat_roughness_y = clampf(v, 0.0f, 1.0f);
        else if (tex_attr.count(a)) ok = b.get_ref(s, Y::kTexture, tex_attr.at(a));
        else if (a == "coat_normal_tex") ok = b.get_ref(s, Y::kTexture, &coat_normal_tex);
        else if (a == "coat_normal_tex_type") {
            std::string t;
            if ((ok = b.get_string(s, &t)) && t != "normal_map" && t != "bump_map") ok = b.fail(s, "invalid normal tex type");
        }
        if (!ok) return false;
    }
    if (!b.named(name)) return false;
    for (int k = 0; k < 3; ++k) {
        m.diffuse_albedo[k] = da[k];
        m.specular_albedo[k] = sa[k];
        m.coat_albedo[k] = ca[k];
    }
    b.y.materials.push_back(m);
    add_name(b.y, name, Y::kMaterial, (uint32_t)b.y.materials.size() - 1);
    return true;
}

bool read_light_material(Block &b)
{
    std::string name;
    float color[3] = {1.0f, 1.0f, 1.0f}, intensity = 1.0f;
    for (size_t i = b.begin + 1; i < b.end; ++i) {
        const Stmt &s = b.st[i];
        if (s.left == "name") {
            if (!b.get_name(s, &name)) return false;
        } else if (s.left == "light_color") {
            if (!b.get_vec(s, color)) return false;
            for (float &x : color) x = maxf(x, 0.0f);
        } else if (s.left == "intensity") {
            if (!b.get_float(s, &intensity)) return false;
            intensity = maxf(intensity, 0.0f);
        }
    }
    if (!b.named(name)) return false;
    ymrt_material_desc m{};
    m.type = YMRT_MATERIAL_LIGHT;
    uint32_t *texs[] = {&m.diffuse_albedo_tex,    &m.specular_albedo_tex,  &m.roughness_x_tex,
                        &m.roughness_y_tex,       &m.specular_weight_tex,  &m.metalness_tex,
                        &m.transmission_weight_tex, &m.coat_albedo_tex,    &m.coat_weight_tex,
                        &m.coat_thickness_tex,    &m.coat_roughness_x_tex, &m.coat_roughness_y_tex,
                        &m.normal_mapping_tex,    &m.bump_mapping_tex};
    for (uint32_t *t : texs) *t = YMRT_INVALID_INDEX;
    for (int k = 0; k < 3; ++k) m.light_color[k] = color[k];
    m.intensity = intensity;
    b.y.materials.push_back(m);
    add_name(b.y, name, Y::kMaterial, (uint32_t)b.y.materials.size() - 1);
    return true;
}

// get_texture_folder_from_scene_file (YMParser.cpp:20-30): <dir>/texture_<stem>_<ext>
std::string texture_folder(const std::string &scene_path)
{
    const size_t slash = scene_path.find_last_of('/');
    const std::string dir = slash == std::string::npos ? "" : scene_path.substr(0, slash + 1);
    const std::string file = slash == std::string::npos ? scene_path : scene_path.substr(slash + 1);
    const size_t dot = file.find_last_of('.');
    const std::string stem = dot == std::string::npos ? file : file.substr(0, dot);
    const std::string ext = dot == std::string::npos ? "" : file.substr(dot + 1);
    return dir + "texture_" + stem + "_" + ext;
}

bool read_image_texture(Block &b)
{
    std::string name, file, cs;
    for (size_t i = b.begin + 1; i < b.end; ++i) {
        const Stmt &s = b.st[i];
        bool ok = true;
        if (s.left == "name") ok = b.get_name(s, &name);
        else if (s.left == "texture_file_name") ok = b.get_string(s, &file);
        else if (s.left == "color_space") ok = b.get_string(s, &cs);
        if (!ok) return false;
    }
    if (!b.named(name)) return false;
    if (file.empty()) {
        b.err = "ImageTexture \"" + name + "\" requires a file";
        return false;
    }
    for (char &c : file)
        if (c == '\\') c = '/';
    ymrt_texture_desc t = texture_base(YMRT_TEXTURE_IMAGE);
    t.wrap_mode = YMRT_WRAP_REPEAT;
    t.color_space = (cs.empty() || cs == "sRGB") ? YMRT_COLOR_SRGB : YMRT_COLOR_LINEAR;
    std::vector<uint8_t> px;
    const std::string path = texture_folder(b.y.scene_path) + "/" + file;
    const std::string e = decode_png(path, &t.width, &t.height, &t.channels, &px);
    if (!e.empty()) {
        // CreateImageTexture returns EMPTY_UINT32: the name stays declared and
        // every reference to it reads "no texture"
        b.y.warnings.push_back("ImageTexture \"" + name + "\": " + e);
        add_name(b.y, name, Y::kTexture, YMRT_INVALID_INDEX);
        return true;
    }
    b.y.texels.push_back(std::move(px));
    b.y.textures.push_back(t);
    add_name(b.y, name, Y::kTexture, (uint32_t)b.y.textures.size() - 1);
    return true;
}

// the procedural and constant texture blocks share one reader: each
// attribute name maps to one field, with the reader's default and clamp
bool read_texture(Block &b)
{
    const std::string &ty = b.type;
    ymrt_texture_desc t;
    if (ty == "ConstantTextureFloat") {
        t = texture_base(YMRT_TEXTURE_CONSTANT_FLOAT);
        t.value[0] = 1.0f;
    } else if (ty == "ConstantTextureRgb") {
        t = texture_base(YMRT_TEXTURE_CONSTANT_RGB);
        t.value[0] = t.value[1] = t.value[2] = 0.5f;
    } else if (ty == "CheckerBoardTexture") {
        t = texture_base(YMRT_TEXTURE_CHECKERBOARD);
        t.frequency = 4.0f;
    } else if (ty == "NoiseTexture") {
        t = texture_base(YMRT_TEXTURE_NOISE);
        t.frequency = 1.0f;
        t.normalized = 1;
    } else if (ty == "NoiseTextureFBM" || ty == "NoiseTextureTurbulence") {
        t = texture_base(ty == "NoiseTextureFBM" ? YMRT_TEXTURE_FBM : YMRT_TEXTURE_TURBULENCE);
        t.frequency = 16.0f;
        t.lacunarity = 2.0f;
        t.gain = 0.5f;
        t.layer_count = 4;
        t.amplitude = 1.0f;
    } else if (ty == "NoiseTextureMarble") {
        t = texture_base(YMRT_TEXTURE_MARBLE);
        t.frequency = 16.0f;
        t.lacunarity = 2.0f;
        t.gain = 0.5f;
        t.layer_count = 4;
        t.variation = 10.0f;
        t.turbulence_axes = 7;
    } else if (ty == "NoiseTextureWood") {
        t = texture_base(YMRT_TEXTURE_WOOD);
        t.frequency = 2.0f;
        t.variation = 10.0f;
    } else if (ty == "NoiseTexturePolkaDot") {
        t = texture_base(YMRT_TEXTURE_POLKA_DOT);
        t.frequency = 8.0f;
        t.radius = 0.35f;
    } else {  // NoiseTextureWave
        t = texture_base(YMRT_TEXTURE_WAVE);
        t.frequency = 0.1f;
        t.lacunarity = 2.0f;
        t.gain = 0.5f;
        t.layer_count = 2;
        t.frequency1 = 2.0f;
        t.lacunarity1 = 2.0f;
        t.gain1 = 0.5f;
        t.layer_count1 = 4;
    }
    const bool two_inputs = t.type >= YMRT_TEXTURE_CHECKERBOARD;
    // frequency is clamped to >= 0.001 by every reader but the marble,
    // polka-dot and wave readers; lacunarity is an accepted attribute of the
    // fBm, turbulence and marble blocks that their readers never read
    const bool clamp_freq = t.type == YMRT_TEXTURE_CHECKERBOARD || t.type == YMRT_TEXTURE_NOISE ||
                            t.type == YMRT_TEXTURE_FBM || t.type == YMRT_TEXTURE_TURBULENCE ||
                            t.type == YMRT_TEXTURE_WOOD;
    std::string name;
    for (size_t i = b.begin + 1; i < b.end; ++i) {
        const Stmt &s = b.st[i];
        const std::string &a = s.left;
        bool ok = true;
        float v;
        int n;
        if (a == "name") ok = b.get_name(s, &name);
        else if (a == "value") ok = b.get_float(s, &t.value[0]);
        else if (a == "color") ok = b.get_vec(s, t.value);
        else if (a == "texture_black_idx") ok = b.get_ref(s, Y::kTexture, &t.child[0]);
        else if (a == "texture_white_idx") ok = b.get_ref(s, Y::kTexture, &t.child[1]);
        else if (a == "frequency" || a == "freq_0") {
            if ((ok = b.get_float(s, &v))) t.frequency = clamp_freq ? maxf(0.001f, v) : v;
        } else if (a == "normalized") t.normalized = Block::get_bool(s) ? 1 : 0;
        else if (a == "gain" || a == "gain_0") ok = b.get_float(s, &t.gain);
        else if (a == "layer_count" || a == "layer_count_0") {
            if ((ok = b.get_int(s, &n))) t.layer_count = n;
        } else if (a == "amplitude") ok = b.get_float(s, &t.amplitude);
        else if (a == "offset") ok = b.get_float(s, &t.offset);
        else if (a == "variation") ok = b.get_float(s, &t.variation);
        else if (a == "radius") ok = b.get_float(s, &t.radius);
        else if (a == "x_turb" || a == "y_turb" || a == "z_turb") {
            const int bit = a[0] == 'x' ? 1 : (a[0] == 'y' ? 2 : 4);
            t.turbulence_axes = Block::get_bool(s) ? (t.turbulence_axes | bit) : (t.turbulence_axes & ~bit);
        } else if (a == "lacunarity_0") ok = b.get_float(s, &t.lacunarity);
        else if (a == "freq_1") ok = b.get_float(s, &t.frequency1);
        else if (a == "lacunarity_1") ok = b.get_float(s, &t.lacunarity1);
        else if (a == "gain_1") ok = b.get_float(s, &t.gain1);
        else if (a == "layer_count_1") {
            if ((ok = b.get_int(s, &n))) t.layer_count1 = n;
        }
        if (!ok) return false;
    }
    if (!b.named(name)) return false;
    if (two_inputs && (t.child[0] == YMRT_INVALID_INDEX || t.child[1] == YMRT_INVALID_INDEX)) {
        b.err = ty + " \"" + name + "\" requires both input textures";
        return false;
    }
    b.y.textures.push_back(t);
    b.y.texels.emplace_back();
    add_name(b.y, name, Y::kTexture, (uint32_t)b.y.textures.size() - 1);
    return true;
}

bool read_distant_light(Block &b)
{
    std::string name;
    float color[3] = {1.0f, 1.0f, 1.0f}, intensity = 1.0f, theta = 5.0f;
    uint32_t xf = YMRT_INVALID_INDEX;
    for (size_t i = b.begin + 1; i < b.end; ++i) {
        const Stmt &s = b.st[i];
        bool ok = true;
        if (s.left == "name") ok = b.get_name(s, &name);
        else if (s.left == "light_color") {
            ok = b.get_vec(s, color);
            for (float &x : color) x = maxf(x, 0.0f);
        } else if (s.left == "transform_index") ok = b.get_ref(s, Y::kTransform, &xf);
        else if (s.left == "intensity" && (ok = b.get_float(s, &intensity))) intensity = maxf(0.0f, intensity);
        else if (s.left == "theta_max" && (ok = b.get_float(s, &theta))) theta = maxf(0.0f, theta);
        if (!ok) return false;
    }
    if (!b.named(name)) return false;
    if (xf == YMRT_INVALID_INDEX) {
        b.err = "DistantLight \"" + name + "\" requires a transform";
        return false;
    }
    ymrt_distant_light_desc d{};
    for (int k = 0; k < 3; ++k) d.light_color[k] = color[k];
    d.intensity = intensity;
    d.theta_max = theta;
    std::memcpy(d.object_to_world, b.y.transforms[xf].c, 64);
    std::memcpy(d.world_to_object, b.y.inverses[xf].c, 64);
    b.y.distant.push_back(d);
    add_name(b.y, name, Y::kDistant, (uint32_t)b.y.distant.size() - 1);
    return true;
}

bool read_volume(Block &b)
{
    std::string name;
    float sigma[3] = {0.05f, 0.05f, 0.05f}, albedo[3] = {1.0f, 1.0f, 1.0f}, g = 0.0f;
    uint32_t xf = YMRT_INVALID_INDEX, density = YMRT_INVALID_INDEX;
    for (size_t i = b.begin + 1; i < b.end; ++i) {
        const Stmt &s = b.st[i];
        bool ok = true;
        if (s.left == "name") ok = b.get_name(s, &name);
        else if (s.left == "transform_index") ok = b.get_ref(s, Y::kTransform, &xf);
        else if (s.left == "sigma_t") {
            ok = b.get_vec(s, sigma);
            for (float &x : sigma) x = maxf(0.0f, x);
        } else if (s.left == "albedo") {
            ok = b.get_vec(s, albedo);
            for (float &x : albedo) x = maxf(0.0f, x);
        } else if (s.left == "g" && (ok = b.get_float(s, &g))) g = clampf(g, 0.0f, 1.0f);
        else if (s.left == "density_texture_index") ok = b.get_ref(s, Y::kTexture, &density);
        if (!ok) return false;
    }
    if (!b.named(name)) return false;
    if (xf == YMRT_INVALID_INDEX) {
        b.err = "Volume \"" + name + "\" requires a transform";
        return false;
    }
    ymrt_volume_desc v{};
    // the texture's index, or none (-1) for an image whose file failed to load
    v.density_texture = density == YMRT_INVALID_INDEX ? -1 : (int32_t)density;
    for (int k = 0; k < 3; ++k) {
        v.sigma_t[k] = sigma[k];
        v.albedo[k] = albedo[k];
    }
    v.g = g;
    v.density_grid = -1;
    std::memcpy(v.world_to_object, b.y.inverses[xf].c, 64);
    b.y.volumes.push_back(v);
    add_name(b.y, name, Y::kVolume, (uint32_t)b.y.volumes.size() - 1);
    return true;
}

bool read_block(Block &b)
{
    const std::string &t = b.type;
    if (t == "Camera") return read_camera(b);
    if (t == "Transform") return read_transform(b);
    if (t == "Mesh") return read_mesh(b);
    if (t == "Cube") return read_cube(b);
    if (t == "Sphere") return read_sphere(b);
    if (t == "PrimitiveInstance") return read_instance(b);
    if (t == "DefaultMaterial") return read_default_material(b);
    if (t == "LightMaterial") return read_light_material(b);
    if (t == "ImageTexture") return read_image_texture(b);
    if (t == "DistantLight") return read_distant_light(b);
    if (t == "Volume") return read_volume(b);
    return read_texture(b);
}

}  // namespace

void ymrt_ymtmp::finish()
{
    mesh_descs.clear();
    for (const Mesh &m : meshes) {
        ymrt_mesh_desc d{};
        d.triangle_count = (uint32_t)(m.pidx.size() / 3);
        d.position_count = (uint32_t)(m.pos.size() / 3);
        d.position_indices = m.pidx.data();
        d.positions = m.pos.data();
        d.normal_count = (uint32_t)(m.nrm.size() / 3);
        d.normal_indices = m.nidx.data();
        d.normals = m.nrm.data();
        d.texcoord_count = (uint32_t)(m.uv.size() / 2);
        d.texcoord_indices = m.tidx.empty() ? nullptr : m.tidx.data();
        d.texcoords = m.uv.empty() ? nullptr : m.uv.data();
        mesh_descs.push_back(d);
    }
    for (size_t i = 0; i < textures.size(); ++i) textures[i].texels = texels[i].empty() ? nullptr : texels[i].data();
    desc = ymrt_scene_desc{};
    desc.mesh_count = (uint32_t)mesh_descs.size();
    desc.meshes = mesh_descs.data();
    desc.sphere_count = (uint32_t)spheres.size();
    desc.spheres = spheres.data();
    desc.instance_count = (uint32_t)instances.size();
    desc.instances = instances.data();
    desc.material_count = (uint32_t)materials.size();
    desc.materials = materials.data();
    desc.texture_count = (uint32_t)textures.size();
    desc.textures = textures.data();
    desc.volume_count = (uint32_t)volumes.size();
    desc.volumes = volumes.data();
    desc.distant_light_count = (uint32_t)distant.size();
    desc.distant_lights = distant.data();
    desc.camera = camera;
}

namespace {

int parse_into(ymrt_ymtmp *y, const std::string &text)
{
    // per line: keep the statement alphabet, drop '#' lines, join
    std::string content;
    {
        std::istringstream in(text);
        std::string line;
        while (std::getline(in, line)) {
            std::string kept;
            for (char c : line)
                if (keep_char(c)) kept.push_back(c);
            if (kept.empty() || kept[0] == '#') continue;
            content += kept;
        }
    }
    std::vector<std::string> texts;
    size_t start = 0;
    for (size_t end = 0; end < content.size(); ++end) {
        if (content[end] != ';') continue;
        if (end > start) texts.push_back(content.substr(start, end - start));
        start = end + 1;
    }
    if (start < content.size()) texts.push_back(content.substr(start));
    std::vector<Stmt> st;
    for (size_t i = 0; i < texts.size(); ++i) {
        if (texts[i][0] == '#') continue;
        Stmt s;
        YM_REQUIRE(parse_statement(texts[i], i, &s), "illegal statement " + std::to_string(i) + ": " + texts[i]);
        st.push_back(s);
    }
    // check_statements_valid (765-814)
    const auto &sets = attribute_sets();
    std::string open;
    for (const Stmt &s : st) {
        if (s.bracket) {
            if (s.left == "Begin") {
                YM_REQUIRE(open.empty(), "nested Begin(" + s.right + ") at statement " + std::to_string(s.line));
                open = s.right;
            } else {
                YM_REQUIRE(!open.empty() && open == s.right,
                           "End(" + s.right + ") does not match at statement " + std::to_string(s.line));
                open.clear();
            }
        } else {
            YM_REQUIRE(!open.empty(), "attribute " + s.left + " outside a block at statement " + std::to_string(s.line));
            const auto &v = sets.at(open);
            YM_REQUIRE(std::find(v.begin(), v.end(), s.left) != v.end(),
                       "unknown attribute " + s.left + " of " + open + " at statement " + std::to_string(s.line));
        }
    }
    YM_REQUIRE(open.empty(), "Begin(" + open + ") is not closed");
    // load_scene_asset (2280-2328)
    for (size_t i = 0; i < st.size(); ++i) {
        if (!(st[i].bracket && st[i].left == "Begin")) continue;
        size_t e = i + 1;
        while (e < st.size() && !(st[e].bracket && st[e].left == "End")) ++e;
        Block b{*y, st, i, e, st[i].right, ""};
        try {
            if (!read_block(b)) y->warnings.push_back(b.err);
        } catch (const std::runtime_error &ex) {
            ym::set_error(ex.what());
            return YMRT_ERR_UNSUPPORTED;
        }
        i = e;
    }
    y->finish();
    return YMRT_OK;
}

}  // namespace

extern "C" {

int ymrt_ymtmp_parse(const char *text, size_t length, const char *scene_path, ymrt_ymtmp **out)
{
    YM_REQUIRE(out != nullptr && (text != nullptr || length == 0), "NULL argument");
    *out = nullptr;
    auto *y = new (std::nothrow) ymrt_ymtmp();
    if (!y) {
        ym::set_error("out of host memory");
        return YMRT_ERR_OUT_OF_MEMORY;
    }
    y->scene_path = scene_path ? scene_path : "";
    for (char &c : y->scene_path)
        if (c == '\\') c = '/';
    y->camera.position[2] = 10.0f;  // the Camera defaults when the file has no Camera block
    y->camera.w[2] = 6.0f;
    y->camera.fov = radians(45.0f);
    y->camera.aspect_ratio = 1.0f;
    int rc;
    try {
        rc = parse_into(y, std::string(text ? text : "", length));
    } catch (const std::exception &e) {
        ym::set_error(std::string("ymtmp: ") + e.what());
        rc = YMRT_ERR_INTERNAL;
    }
    if (rc != YMRT_OK) {
        delete y;
        return rc;
    }
    *out = y;
    return YMRT_OK;
}

int ymrt_ymtmp_load(const char *path, ymrt_ymtmp **out)
{
    YM_REQUIRE(path != nullptr && out != nullptr, "NULL argument");
    std::string p(path);
    for (char &c : p)
        if (c == '\\') c = '/';
    YM_REQUIRE(p.size() >= 6 && p.compare(p.size() - 6, 6, ".ymtmp") == 0, "invalid file type (expected .ymtmp)");
    std::ifstream f(p, std::ios::binary);
    YM_REQUIRE(f.good(), "failed to open " + p);
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string text = ss.str();
    return ymrt_ymtmp_parse(text.data(), text.size(), p.c_str(), out);
}

const ymrt_scene_desc *ymrt_ymtmp_scene(const ymrt_ymtmp *y) { return y ? &y->desc : nullptr; }

uint32_t ymrt_ymtmp_warning_count(const ymrt_ymtmp *y) { return y ? (uint32_t)y->warnings.size() : 0u; }

const char *ymrt_ymtmp_warning(const ymrt_ymtmp *y, uint32_t i)
{
    return (y && i < y->warnings.size()) ? y->warnings[i].c_str() : nullptr;
}

int ymrt_ymtmp_find(const ymrt_ymtmp *y, const char *name, uint32_t *kind, uint32_t *index)
{
    YM_REQUIRE(y != nullptr && name != nullptr && kind != nullptr && index != nullptr, "NULL argument");
    const auto it = y->names.find(name);
    YM_REQUIRE(it != y->names.end(), std::string("no asset named ") + name);
    static const uint32_t kinds[] = {YMRT_YMTMP_TRANSFORM, YMRT_YMTMP_MESH,     YMRT_YMTMP_SPHERE,
                                     YMRT_YMTMP_MATERIAL,  YMRT_YMTMP_TEXTURE,  YMRT_YMTMP_INSTANCE,
                                     YMRT_YMTMP_DISTANT_LIGHT, YMRT_YMTMP_VOLUME};
    *kind = kinds[it->second.kind];
    *index = it->second.index;
    return YMRT_OK;
}

int ymrt_ymtmp_transform(const ymrt_ymtmp *y, uint32_t index, float *object_to_world, float *world_to_object)
{
    YM_REQUIRE(y != nullptr, "NULL argument");
    YM_REQUIRE(index < y->transforms.size(), "transform index out of range");
    if (object_to_world) std::memcpy(object_to_world, y->transforms[index].c, 64);
    if (world_to_object) std::memcpy(world_to_object, y->inverses[index].c, 64);
    return YMRT_OK;
}

int ymrt_ymtmp_free(ymrt_ymtmp *y)
{
    delete y;
    return YMRT_OK;
}

}  // extern "C"
