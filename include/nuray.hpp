// nuray.hpp -- C++ facade over the C ABI (nuray.h) with the reference's host API.
//
// Mirrors, without Qt:
//   Config    src/utils/config.h:7-22      (k=v items, typed getters with defaults)
//   Camera    src/scene/camera.h:8-34      (Camera(fov, asp, pos, yaw, pitch, roll))
//   Loader    src/scene/loader.h:8-19      (scene description, OBJ + MTL, loader.cpp:26-361)
//   Texture   src/scene/texture.cpp:5-31   (8-bit image -> pow(c / 255, 2.2); PNG / BMP / PPM here,
//                                          link with -lz)
//   Renderer  src/renderer/renderer.h:15-35 (prepare / render / setEnvmap) on the GPU ptnee path
//   Image     QImage RGB888 output, saved as BMP (cmd.cpp:112 saves result.bmp)
// plus RenderFunction, a torch.autograd.Function-style pair forward(ctx, theta) -> image and
// backward(ctx, dL/dimage) -> dL/dtheta over nr_render_ptnee / nr_adjoint_ptnee.
//
// Everything throws nuray::Error on failure (the C ABI's nr_last_error() text).
#ifndef NURAY_HPP
#define NURAY_HPP

#include <zlib.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "nuray.h"

namespace nuray {

struct Error : std::runtime_error {
    explicit Error(const std::string& m) : std::runtime_error(m) {}
};

inline void check(int rc) {
    if (rc < 0) throw Error(nr_last_error());
}

struct vec3 {
    float v[3] = {0.0f, 0.0f, 0.0f};
    vec3() = default;
    vec3(float a) : v{a, a, a} {}
    vec3(float a, float b, float c) : v{a, b, c} {}
    float& operator[](int i) { return v[i]; }
    float operator[](int i) const { return v[i]; }
    float norm2() const { return v[0] * v[0] + v[1] * v[1] + v[2] * v[2]; }
};

// ------------------------------------------------------------------------------------- Config
// Config (config.cpp:12-104): whitespace-separated k=v items; a getter returns its default when
// the key is missing or does not parse.  getValueStr returns "" for a missing key, as the
// reference does (config.cpp:22-30), whatever the default.
class Config {
    std::map<std::string, std::string> mp_;

public:
    Config() = default;
    explicit Config(const std::string& s) { fromString(s); }
    void fromString(const std::string& s) {
        std::stringstream ss(s);
        std::string kv;
        while (ss >> kv) addItem(kv);
    }
    void addItem(const std::string& kv) {
        const auto p = kv.find('=');
        if (p == std::string::npos) return;
        mp_[kv.substr(0, p)] = kv.substr(p + 1);
    }
    bool has(const std::string& k) const { return mp_.count(k) != 0; }
    std::string getValueStr(const std::string& k, const std::string& /*def*/) const {
        auto it = mp_.find(k);
        return it == mp_.end() ? std::string() : it->second;
    }
    int getValueInt(const std::string& k, int def) const {
        auto it = mp_.find(k);
        if (it == mp_.end()) return def;
        std::stringstream ss(it->second);
        int x;
        return (ss >> x) ? x : def;
    }
    float getValueFloat(const std::string& k, float def) const {
        auto it = mp_.find(k);
        if (it == mp_.end()) return def;
        std::stringstream ss(it->second);
        float x;
        return (ss >> x) ? x : def;
    }
    vec3 getValueVec3(const std::string& k, const vec3& def) const {
        auto it = mp_.find(k);
        if (it == mp_.end()) return def;
        std::string s = it->second;
        for (char& c : s)
            if (c == ',') c = ' ';
        std::stringstream ss(s);
        float x, y, z;
        return (ss >> x >> y >> z) ? vec3(x, y, z) : def;
    }
    void setValueStr(const std::string& k, const std::string& v) { mp_[k] = v; }
    void print(std::ostream& os = std::cout) const {
        for (const auto& kv : mp_) os << kv.first << " = " << kv.second << "\n";
    }
};

// ------------------------------------------------------------------------------------- Camera
struct Camera {
    float fov = 90.0f, asp = 1.0f;
    vec3 pos, euler;  // euler = yaw, pitch, roll in degrees
    Camera() = default;
    Camera(float fov_, float asp_, vec3 p, float yaw, float pitch, float roll)
        : fov(fov_), asp(asp_), pos(p), euler(yaw, pitch, roll) {}
    nr_camera c() const {
        nr_camera r;
        for (int i = 0; i < 3; ++i) { r.pos[i] = pos[i]; r.euler[i] = euler[i]; }
        r.fov = fov;
        r.aspect = asp;
        return r;
    }
};

// ------------------------------------------------------------------------------------- images
// An 8-bit RGB image (first row = top), with the linear mean beside it when rendered.
struct Image {
    int w = 0, h = 0;
    std::vector<uint8_t> rgb;    // w * h * 3
    std::vector<float> linear;   // w * h * 3, mean radiance
    void resize(int W, int H) { w = W; h = H; rgb.assign(size_t(W) * H * 3, 0); linear.assign(size_t(W) * H * 3, 0.0f); }
    // 24-bit uncompressed BMP, bottom-up rows padded to 4 bytes.
    void saveBMP(const std::string& path) const {
        const int row = (w * 3 + 3) & ~3;
        const uint32_t size = 54u + uint32_t(row) * uint32_t(h);
        std::vector<uint8_t> f(size, 0);
        auto put32 = [&](size_t o, uint32_t x) { for (int i = 0; i < 4; ++i) f[o + i] = uint8_t(x >> (8 * i)); };
        auto put16 = [&](size_t o, uint16_t x) { f[o] = uint8_t(x); f[o + 1] = uint8_t(x >> 8); };
        f[0] = 'B'; f[1] = 'M';
        put32(2, size); put32(10, 54); put32(14, 40); put32(18, uint32_t(w)); put32(22, uint32_t(h));
        put16(26, 1); put16(28, 24); put32(34, uint32_t(row) * uint32_t(h)); put32(38, 2835); put32(42, 2835);
        for (int y = 0; y < h; ++y) {
            uint8_t* dst = f.data() + 54 + size_t(h - 1 - y) * row;
            const uint8_t* src = rgb.data() + size_t(y) * w * 3;
            for (int x = 0; x < w; ++x) {
                dst[3 * x + 0] = src[3 * x + 2];
                dst[3 * x + 1] = src[3 * x + 1];
                dst[3 * x + 2] = src[3 * x + 0];
            }
        }
        std::ofstream os(path, std::ios::binary);
        if (!os || !os.write(reinterpret_cast<const char*>(f.data()), f.size())) throw Error("cannot write " + path);
    }
};

// 8-bit RGBA texels (first row = top), as Texture::load keeps them before its 2.2 decode.
struct Texture {
    int w = 0, h = 0;
    std::vector<uint8_t> rgba;
    // PNG, BMP (24 / 32-bit uncompressed) or binary PPM (P6, maxval 255).  JPEG is refused.
    static Texture load(const std::string& path) {
        std::ifstream is(path, std::ios::binary);
        if (!is) throw Error("cannot open texture " + path);
        std::vector<uint8_t> d((std::istreambuf_iterator<char>(is)), std::istreambuf_iterator<char>());
        Texture t;
        auto rd32 = [&](size_t o) { return uint32_t(d[o]) | uint32_t(d[o + 1]) << 8 | uint32_t(d[o + 2]) << 16 | uint32_t(d[o + 3]) << 24; };
        if (d.size() > 54 && d[0] == 'B' && d[1] == 'M') {
            const uint32_t off = rd32(10);
            const int32_t W = int32_t(rd32(18)), Hs = int32_t(rd32(22));
            const int bpp = d[28] | d[29] << 8;
            if (rd32(30) != 0 || (bpp != 24 && bpp != 32) || W <= 0 || Hs == 0)
                throw Error("unsupported BMP (only uncompressed 24 / 32-bit): " + path);
            const int H = Hs < 0 ? -Hs : Hs, bypp = bpp / 8, row = (W * bypp + 3) & ~3;
            if (d.size() < off + size_t(row) * H) throw Error("truncated BMP " + path);
            t.w = W; t.h = H;
            t.rgba.resize(size_t(W) * H * 4);
            for (int y = 0; y < H; ++y) {
                const uint8_t* s = d.data() + off + size_t(Hs < 0 ? y : H - 1 - y) * row;
                for (int x = 0; x < W; ++x) {
                    uint8_t* o = t.rgba.data() + (size_t(y) * W + x) * 4;
                    o[0] = s[bypp * x + 2]; o[1] = s[bypp * x + 1]; o[2] = s[bypp * x + 0];
                    o[3] = bypp == 4 ? s[bypp * x + 3] : 255;
                }
            }
            return t;
        }
        if (d.size() > 2 && d[0] == 'P' && d[1] == '6') {
            std::string hdr(d.begin(), d.begin() + std::min<size_t>(d.size(), 64));
            std::stringstream ss(hdr.substr(2));
            int W, H, mx;
            if (!(ss >> W >> H >> mx) || mx != 255 || W <= 0 || H <= 0) throw Error("unsupported PPM " + path);
            const size_t off = 2 + size_t(ss.tellg()) + 1;
            if (d.size() < off + size_t(W) * H * 3) throw Error("truncated PPM " + path);
            t.w = W; t.h = H;
            t.rgba.resize(size_t(W) * H * 4);
            for (size_t i = 0; i < size_t(W) * H; ++i) {
                for (int c = 0; c < 3; ++c) t.rgba[4 * i + c] = d[off + 3 * i + c];
                t.rgba[4 * i + 3] = 255;
            }
            return t;
        }
        if (d.size() > 8 && std::memcmp(d.data(), "\x89PNG\r\n\x1a\n", 8) == 0) return load_png(d, path);
        if (d.size() > 2 && d[0] == 0xFF && d[1] == 0xD8)
            throw Error("JPEG textures are not supported (convert to PNG, BMP or PPM): " + path);
        throw Error("unsupported texture format (PNG, BMP or binary PPM): " + path);
    }
    // PNG (ISO/IEC 15948): IHDR, PLTE, tRNS and the IDAT stream (inflated by zlib); every scanline's
    // filter (None, Sub, Up, Average, Paeth) undone; greyscale, RGB, palette, grey + alpha and RGBA
    // at 1-16 bits per sample.  16-bit samples keep their high byte.  Interlaced images are refused.
    static Texture load_png(const std::vector<uint8_t>& d, const std::string& path) {
        auto be32 = [&d](size_t o) {
            return uint32_t(d[o]) << 24 | uint32_t(d[o + 1]) << 16 | uint32_t(d[o + 2]) << 8 | uint32_t(d[o + 3]);
        };
        uint32_t W = 0, H = 0;
        int depth = 0, ctype = -1, interlace = 0;
        std::vector<uint8_t> z, plte, trns;
        for (size_t o = 8; o + 12 <= d.size();) {
            const uint32_t len = be32(o);
            if (o + 12 + size_t(len) > d.size()) throw Error("truncated PNG " + path);
            const std::string type(reinterpret_cast<const char*>(&d[o + 4]), 4);
            const uint8_t* c = &d[o + 8];
            if (type == "IHDR" && len >= 13) {
                W = be32(o + 8);
                H = be32(o + 12);
                depth = c[8];
                ctype = c[9];
                interlace = c[12];
            } else if (type == "PLTE") {
                plte.assign(c, c + len);
            } else if (type == "tRNS") {
                trns.assign(c, c + len);
            } else if (type == "IDAT") {
                z.insert(z.end(), c, c + len);
            } else if (type == "IEND") {
                break;
            }
            o += 12 + size_t(len);
        }
        static const int channels[7] = {1, 0, 3, 1, 2, 0, 4};
        if (W == 0 || H == 0 || W > (1u << 15) || H > (1u << 15) || ctype < 0 || ctype > 6 || !channels[ctype] ||
            (depth != 1 && depth != 2 && depth != 4 && depth != 8 && depth != 16) || (ctype == 3 && plte.empty()))
            throw Error("unsupported PNG header: " + path);
        if (interlace) throw Error("interlaced PNG textures are not supported: " + path);
        const int nch = channels[ctype];
        const size_t bits = size_t(nch) * depth, stride = (size_t(W) * bits + 7) / 8, bpp = std::max<size_t>(1, bits / 8);
        std::vector<uint8_t> raw(size_t(H) * (stride + 1));
        z_stream zs{};
        if (inflateInit(&zs) != Z_OK) throw Error("zlib init failed for " + path);
        zs.next_in = z.data();
        zs.avail_in = uInt(z.size());
        zs.next_out = raw.data();
        zs.avail_out = uInt(raw.size());
        const int zr = inflate(&zs, Z_FINISH);
        const size_t got = raw.size() - zs.avail_out;
        inflateEnd(&zs);
        if ((zr != Z_STREAM_END && zr != Z_OK) || got != raw.size()) throw Error("corrupt PNG data in " + path);
        // undo the filters in place, row by row (the previous row is already unfiltered)
        for (uint32_t y = 0; y < H; ++y) {
            uint8_t* row = &raw[size_t(y) * (stride + 1)];
            const int f = row[0];
            uint8_t* cur = row + 1;
            const uint8_t* up = y ? &raw[size_t(y - 1) * (stride + 1) + 1] : nullptr;
            for (size_t i = 0; i < stride; ++i) {
                const int a = i >= bpp ? cur[i - bpp] : 0, b = up ? up[i] : 0, cc = (up && i >= bpp) ? up[i - bpp] : 0;
                int pred = 0;
                switch (f) {
                    case 0: pred = 0; break;
                    case 1: pred = a; break;
                    case 2: pred = b; break;
                    case 3: pred = (a + b) >> 1; break;
                    case 4: {
                        const int p = a + b - cc, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - cc);
                        pred = (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : cc);
                        break;
                    }
                    default: throw Error("bad PNG filter type in " + path);
                }
                cur[i] = uint8_t(cur[i] + pred);
            }
        }
        // one sample (0..255) of channel k of pixel x in an unfiltered row
        auto sample = [&](const uint8_t* row, uint32_t x, int k) -> int {
            if (depth == 8) return row[size_t(x) * nch + k];
            if (depth == 16) return row[(size_t(x) * nch + k) * 2];  // high byte
            const size_t bit = (size_t(x) * nch + k) * depth;
            const int v = (row[bit / 8] >> (8 - depth - int(bit % 8))) & ((1 << depth) - 1);
            return ctype == 3 ? v : v * 255 / ((1 << depth) - 1);
        };
        Texture t;
        t.w = int(W);
        t.h = int(H);
        t.rgba.resize(size_t(W) * H * 4);
        for (uint32_t y = 0; y < H; ++y) {
            const uint8_t* row = &raw[size_t(y) * (stride + 1) + 1];
            for (uint32_t x = 0; x < W; ++x) {
                uint8_t* o = &t.rgba[(size_t(y) * W + x) * 4];
                if (ctype == 3) {
                    const int i = sample(row, x, 0);
                    if (size_t(3 * i + 2) >= plte.size()) throw Error("PNG palette index out of range in " + path);
                    o[0] = plte[3 * i]; o[1] = plte[3 * i + 1]; o[2] = plte[3 * i + 2];
                    o[3] = size_t(i) < trns.size() ? trns[i] : 255;
                } else if (ctype == 0 || ctype == 4) {
                    o[0] = o[1] = o[2] = uint8_t(sample(row, x, 0));
                    o[3] = ctype == 4 ? uint8_t(sample(row, x, 1)) : 255;
                } else {
                    for (int k = 0; k < 3; ++k) o[k] = uint8_t(sample(row, x, k));
                    o[3] = ctype == 6 ? uint8_t(sample(row, x, 3)) : 255;
                }
            }
        }
        return t;
    }
    // Texture::load's decode (texture.cpp:18-20): linear = pow(c / 255, 2.2).
    std::vector<float> linear_rgb() const {
        std::vector<float> out(size_t(w) * h * 3);
        for (size_t i = 0; i < size_t(w) * h; ++i)
            for (int c = 0; c < 3; ++c) out[3 * i + c] = std::pow(float(rgba[4 * i + c]) / 255.0f, 2.2f);
        return out;
    }
};

// Envmap(texture, gain) (envmap.h:8-20).
struct Envmap {
    int w = 0, h = 0;
    std::vector<float> rgb;  // linear
    float gain = 1.0f;
    Envmap() = default;
    Envmap(const Texture& t, float g) : w(t.w), h(t.h), rgb(t.linear_rgb()), gain(g) {}
};

// ------------------------------------------------------------------------------------- scene
// One row of the C ABI's material table (nuray.h): {type, r, g, b, alpha, ior, texture, 0}.
struct Material {
    int type = NR_MAT_LAMBERT;
    vec3 kd;                   // Kd / Ke / F0
    float alpha = 0.5f, ior = 1.0f;  // alpha: GGX roughness, or the Blinn-Phong exponent Ns
    int tex = -1;
    vec3 ks;                   // Blinn-Phong
    int tex_ks = -1;
};

struct Triangle {
    vec3 p[3], t[3], n[3];
    int mat = 0;  // index into Loader::materials()
};

// Loader (loader.cpp:26-361): scene description lines "<file.obj> [-p x y z] [-s scale]
// [-m=light r g b | -m=ggx r g b alpha ior | -m=ggxr r g b alpha | -m=glass r g b ior]",
// OBJ v / vt / vn / f (triangles and quads split (0,1,2), (2,3,0), negative indices relative),
// MTL newmtl / Kd / Ks / Ke / Ns / Ni / map_Kd / map_Ks; Ke > 0 -> light, Ks or map_Ks ->
// Blinn-Phong (Kd, Ks, Ns, map_Kd, map_Ks), else Lambert (map_Kd).  Triangles keep the file order.
class Loader {
    std::vector<Triangle> tris_;
    std::vector<Material> mats_;
    std::vector<Texture> texs_;
    std::map<std::string, int> dict_;
    std::map<std::string, int> tex_dict_;

    static std::string dir_of(const std::string& path) {
        const auto p = path.find_last_of("/\\");
        return p == std::string::npos ? "." : path.substr(0, p);
    }
    int texture(const std::string& path) {
        auto it = tex_dict_.find(path);
        if (it != tex_dict_.end()) return it->second;
        texs_.push_back(Texture::load(path));
        return tex_dict_[path] = int(texs_.size()) - 1;
    }
    static float num(const std::vector<std::string>& b, size_t i) {
        return i < b.size() ? float(std::atof(b[i].c_str())) : 0.0f;
    }

public:
    void fromSceneDescription(const std::string& desc) {
        tris_.clear();
        mats_.clear();
        texs_.clear();
        dict_.clear();
        tex_dict_.clear();
        std::stringstream ds(desc);
        std::string line;
        while (std::getline(ds, line)) {
            std::stringstream ls(line);
            std::string file, op;
            if (!(ls >> file)) continue;
            vec3 pos(0.0f);
            float scale = 1.0f;
            int forcing = -1;
            while (ls >> op) {
                Material m;
                if (op == "-p") {
                    ls >> pos[0] >> pos[1] >> pos[2];
                    continue;
                } else if (op == "-s") {
                    ls >> scale;
                    continue;
                } else if (op == "-m=light") {
                    m.type = NR_MAT_LIGHT;
                    ls >> m.kd[0] >> m.kd[1] >> m.kd[2];
                } else if (op == "-m=ggx") {
                    m.type = NR_MAT_GGX;
                    ls >> m.kd[0] >> m.kd[1] >> m.kd[2] >> m.alpha >> m.ior;
                } else if (op == "-m=ggxr") {
                    m.type = NR_MAT_GGXREFL;
                    ls >> m.kd[0] >> m.kd[1] >> m.kd[2] >> m.alpha;
                } else if (op == "-m=glass") {
                    m.type = NR_MAT_GLASS;
                    ls >> m.kd[0] >> m.kd[1] >> m.kd[2] >> m.ior;
                } else {
                    continue;
                }
                mats_.push_back(m);
                forcing = int(mats_.size()) - 1;
            }
            loadObj(file, pos, scale, forcing);
        }
    }

    void loadMtl(const std::string& file) {
        std::ifstream is(file);
        if (!is) throw Error("cannot open " + file);
        std::string name, line, map_kd, map_ks;
        vec3 kd, ks, ke;
        float ns = 10.0f;
        auto commit = [&]() {
            Material m;
            if (ke.norm2() > 1e-6f) {
                m.type = NR_MAT_LIGHT;
                m.kd = ke;
            } else if (ks.norm2() > 1e-6f || !map_ks.empty()) {
                m.type = NR_MAT_BLINNPHONG;
                m.kd = kd;
                m.ks = ks;
                m.alpha = ns;
                if (!map_kd.empty()) m.tex = texture(map_kd);
                if (!map_ks.empty()) m.tex_ks = texture(map_ks);
            } else {
                m.type = NR_MAT_LAMBERT;
                m.kd = kd;
                if (!map_kd.empty()) m.tex = texture(map_kd);
            }
            mats_.push_back(m);
            dict_[file + ":" + name] = int(mats_.size()) - 1;
            kd = ks = ke = vec3(0.0f);
            ns = 10.0f;
            map_kd.clear();
            map_ks.clear();
        };
        while (std::getline(is, line)) {
            if (!line.empty() && line[0] == '#') continue;
            std::stringstream ss(line);
            std::vector<std::string> b;
            std::string tok;
            while (ss >> tok) b.push_back(tok);
            if (b.empty()) continue;
            if (b[0] == "newmtl") {
                if (!name.empty()) commit();
                name = b.size() > 1 ? b[1] : "";
            } else if (b[0] == "Kd") {
                kd = vec3(num(b, 1), num(b, 2), num(b, 3));
            } else if (b[0] == "Ks") {
                ks = vec3(num(b, 1), num(b, 2), num(b, 3));
            } else if (b[0] == "Ke") {
                ke = vec3(num(b, 1), num(b, 2), num(b, 3));
            } else if (b[0] == "Ns") {
                ns = num(b, 1);
            } else if (b[0] == "map_Kd" && b.size() > 1) {
                map_kd = dir_of(file) + "/" + b[1];
            } else if (b[0] == "map_Ks" && b.size() > 1) {
                map_ks = dir_of(file) + "/" + b[1];
            }
        }
        if (!name.empty()) commit();
    }

    void loadObj(const std::string& file, const vec3& pos, float scale, int forcing = -1) {
        std::ifstream is(file);
        if (!is) throw Error("cannot open " + file);
        std::vector<vec3> v(1), vt(1), vn(1);
        std::string line, mtllib;
        int cur = forcing;
        while (std::getline(is, line)) {
            std::stringstream ss(line);
            std::vector<std::string> b;
            std::string tok;
            while (ss >> tok) b.push_back(tok);
            if (b.empty()) continue;
            if (b[0] == "mtllib" && b.size() > 1) {
                mtllib = dir_of(file) + "/" + b[1];
                if (forcing < 0) loadMtl(mtllib);
            } else if (b[0] == "usemtl" && b.size() > 1) {
                if (forcing < 0) {
                    auto it = dict_.find(mtllib + ":" + b[1]);
                    if (it == dict_.end()) throw Error("cannot find material " + mtllib + ":" + b[1]);
                    cur = it->second;
                }
            } else if (b[0] == "v") {
                // v * scale + position (loader.cpp:274), in float as the reference's vec3 ops
                v.push_back(vec3(num(b, 1) * scale + pos[0], num(b, 2) * scale + pos[1], num(b, 3) * scale + pos[2]));
            } else if (b[0] == "vt") {
                vt.push_back(vec3(num(b, 1), num(b, 2), 0.0f));
            } else if (b[0] == "vn") {
                vn.push_back(vec3(num(b, 1), num(b, 2), num(b, 3)));
            } else if (b[0] == "f") {
                const size_t k = b.size() - 1;
                if (k != 3 && k != 4) throw Error("unsupported face with " + std::to_string(k) + " vertices in " + file);
                std::vector<int> iv(k), it(k), in(k);
                for (size_t i = 0; i < k; ++i) {
                    const std::string& s = b[i + 1];
                    int idx[3] = {0, 0, 0};
                    size_t start = 0;
                    for (int j = 0; j < 3; ++j) {
                        const size_t slash = s.find('/', start);
                        const size_t end = slash == std::string::npos ? s.size() : slash;
                        if (start < end) idx[j] = std::atoi(s.substr(start, end - start).c_str());
                        if (slash == std::string::npos) { start = s.size(); } else { start = slash + 1; }
                    }
                    iv[i] = idx[0] < 0 ? idx[0] + int(v.size()) : idx[0];
                    it[i] = idx[1] < 0 ? idx[1] + int(vt.size()) : idx[1];
                    in[i] = idx[2] < 0 ? idx[2] + int(vn.size()) : idx[2];
                    if (iv[i] <= 0 || iv[i] >= int(v.size())) throw Error("vertex index out of range in " + file);
                }
                if (cur < 0) throw Error("face without a material in " + file);
                auto emit = [&](int a, int bb, int c) {
                    Triangle t;
                    const int ids[3] = {a, bb, c};
                    for (int q = 0; q < 3; ++q) {
                        const int j = ids[q];
                        t.p[q] = v[iv[j]];
                        t.t[q] = (it[j] > 0 && it[j] < int(vt.size())) ? vt[it[j]] : vec3(0.0f);
                        t.n[q] = (in[j] > 0 && in[j] < int(vn.size())) ? vn[in[j]] : vec3(0.0f);
                    }
                    t.mat = cur;
                    tris_.push_back(t);
                };
                emit(0, 1, 2);
                if (k == 4) emit(2, 3, 0);
            }
        }
    }

    const std::vector<Triangle>& getTriangles() const { return tris_; }
    const std::vector<Material>& materials() const { return mats_; }
    const std::vector<Texture>& textures() const { return texs_; }
};

// ------------------------------------------------------------------------------------- Renderer
// Renderer (renderer.h:15-35) for "ptnee", "pt" and "nrc".  render() reads the Config keys imgw, imgh,
// spp, prr, mis (rendererptls.cpp), seed, maxdepth, rrdepth, and for "nrc" frames (frames of
// nrc_spp samples each, training online), nrc_spread, nrc_train_paths.
class Renderer {
    nr_scene* s_ = nullptr;
    nr_nrc* cache_ = nullptr;
    std::string name_;
    Envmap env_;
    bool has_env_ = false;

public:
    explicit Renderer(const std::string& name = "ptnee") : name_(name) {
        if (name_ != "ptnee" && name_ != "pt" && name_ != "nrc")
            throw Error("renderer '" + name_ + "' is not available (ptnee, pt, nrc)");
        check(nr_init(0));
    }
    Renderer(const Renderer&) = delete;
    Renderer& operator=(const Renderer&) = delete;
    ~Renderer() {
        if (s_) nr_scene_destroy(s_);
        if (cache_) nr_nrc_destroy(cache_);
    }
    nr_scene* handle() const { return s_; }
    const std::string& name() const { return name_; }

    void setEnvmap(const Envmap& e) {
        env_ = e;
        has_env_ = true;
        if (s_) check(nr_scene_set_envmap(s_, env_.w, env_.h, env_.rgb.data(), env_.gain));
    }

    // Renderer::prepare (renderer.cpp:30-39): upload triangles, materials and textures, build the BVH.
    void prepare(const std::vector<Triangle>& tris, const std::vector<Material>& mats,
                 const std::vector<Texture>& texs = {}) {
        const size_t N = tris.size();
        std::vector<float> p(N * 9), n(N * 9), uv(N * 6), m(mats.size() * 8);
        std::vector<int32_t> id(N);
        for (size_t i = 0; i < N; ++i) {
            for (int k = 0; k < 3; ++k)
                for (int c = 0; c < 3; ++c) {
                    p[i * 9 + 3 * k + c] = tris[i].p[k][c];
                    n[i * 9 + 3 * k + c] = tris[i].n[k][c];
                }
            for (int k = 0; k < 3; ++k) {
                uv[i * 6 + 2 * k] = tris[i].t[k][0];
                uv[i * 6 + 2 * k + 1] = tris[i].t[k][1];
            }
            id[i] = tris[i].mat;
        }
        for (size_t j = 0; j < mats.size(); ++j) {
            const Material& q = mats[j];
            const float row[8] = {float(q.type), q.kd[0], q.kd[1], q.kd[2], q.alpha, q.ior, float(q.tex), 0.0f};
            std::memcpy(&m[j * 8], row, sizeof(row));
        }
        if (s_) nr_scene_destroy(s_);
        s_ = nr_scene_create(p.data(), n.data(), uv.data(), id.data(), uint32_t(N), m.data(), uint32_t(mats.size()));
        if (!s_) throw Error(nr_last_error());
        for (size_t k = 0; k < texs.size(); ++k)
            check(nr_scene_set_texture_rgba8(s_, int32_t(k), texs[k].w, texs[k].h, texs[k].rgba.data()));
        for (size_t j = 0; j < mats.size(); ++j)
            if (mats[j].type == NR_MAT_BLINNPHONG) {
                const float ext[4] = {mats[j].ks[0], mats[j].ks[1], mats[j].ks[2], float(mats[j].tex_ks)};
                check(nr_scene_set_material_ext(s_, int32_t(j), ext));
            }
        if (has_env_) check(nr_scene_set_envmap(s_, env_.w, env_.h, env_.rgb.data(), env_.gain));
        check(nr_bvh_build_sah(s_, 0));
    }
    void prepare(const Loader& l) { prepare(l.getTriangles(), l.materials(), l.textures()); }

    static nr_params params_of(const Config& cfg) {
        nr_params p{};
        p.width = cfg.getValueInt("imgw", 640);
        p.height = cfg.getValueInt("imgh", 480);
        p.spp = cfg.getValueInt("spp", 4);
        p.prr = cfg.getValueFloat("prr", 0.8f);
        p.mis = cfg.getValueInt("mis", 1);
        p.seed = uint32_t(cfg.getValueInt("seed", 0));
        p.maxdepth = cfg.getValueInt("maxdepth", 0);
        p.rrdepth = cfg.getValueInt("rrdepth", 0);
        p.env_is = cfg.getValueInt("envmap_is", 0);
        p.path_kernel = cfg.getValueInt("path_kernel", 2);  // automatic unless the Config says otherwise
        return p;
    }

    // Renderer::render (renderer.h:22-29): blocks; progress(fraction) after each wavefront batch;
    // a zero `con_flag` (set from any thread) cancels, as the reference's con_flag does: a watcher
    // thread raises the render's cancel word, which the render reads between bounces and batches.
    void render(const Camera& cam, Image& img, const Config& cfg, std::function<void(float)> progress = {},
                std::atomic<int>* con_flag = nullptr) {
        if (!s_) throw Error("render before prepare");
        nr_params p = params_of(cfg);
        img.resize(p.width, p.height);
        const nr_camera c = cam.c();
        if (name_ == "ptnee" || name_ == "pt") {
            p.integrator = name_ == "pt" ? 1 : 0;
            struct Ctx { std::function<void(float)>* f; std::atomic<int>* flag; volatile int32_t* cancel; };
            volatile int32_t cancel = 0;
            Ctx ctx{&progress, con_flag, &cancel};
            auto cb = [](float x, void* u) {
                Ctx* k = static_cast<Ctx*>(u);
                if (*k->f) (*k->f)(x);
                if (k->flag && k->flag->load() == 0) *k->cancel = 1;
            };
            std::atomic<bool> done{false};
            std::thread watcher;
            if (con_flag)
                watcher = std::thread([&] {
                    while (!done.load()) {
                        if (con_flag->load() == 0) {
                            cancel = 1;
                            return;
                        }
                        std::this_thread::sleep_for(std::chrono::milliseconds(2));
                    }
                });
            const int rc = nr_render_ptnee(s_, &c, &p, img.linear.data(), nullptr, nullptr, img.rgb.data(), cb, &ctx,
                                           &cancel);
            done.store(true);
            if (watcher.joinable()) watcher.join();
            check(rc);
            if (rc == NR_CANCELLED) nr_render_fetch(s_, img.linear.data(), nullptr, nullptr, img.rgb.data());
            return;
        }
        // nrc: `frames` frames of nrc_spp samples per pixel, the cache training online
        if (!cache_) {
            cache_ = nr_nrc_create(uint32_t(cfg.getValueInt("seed", 0)), cfg.getValueFloat("nrc_lr", 1e-3f), 0);
            if (!cache_) throw Error(nr_last_error());
        }
        const int frames = cfg.getValueInt("frames", std::max(1, p.spp));
        nr_nrc_params np{cfg.getValueInt("nrc_train_paths", 65536), cfg.getValueInt("nrc_train_maxlen", 5),
                         cfg.getValueFloat("nrc_spread", 0.01f), 1, 1, cfg.getValueInt("nrc_train_steps", 4),
                         cfg.getValueInt("nrc_min_depth", 1)};
        const int spp = cfg.getValueInt("nrc_spp", 1);
        for (int f = 0; f < frames; ++f) {
            if (con_flag && con_flag->load() == 0) break;
            nr_params q = p;
            q.spp = spp;
            q.spp_offset = f * spp;
            q.accumulate = f > 0;
            check(nr_render_nrc(s_, cache_, &c, &q, &np, nullptr));
            if (progress) progress(float(f + 1) / float(frames));
        }
        check(nr_render_fetch(s_, img.linear.data(), nullptr, nullptr, img.rgb.data()));
    }
};

// ------------------------------------------------------------------------------------- autograd
// torch.autograd.Function-style differentiable render: forward saves what backward needs in the
// context; backward maps dL/dimage (W*H*3 floats) to dL/dtheta (materials x {r, g, b, alpha}).
// The render runs as one wavefront batch (images up to 64 M samples).  The context keeps theta, so
// a backward after other forwards on the same renderer (other views, other parameters) puts its own
// parameters back on the scene first; the adjoint then re-renders its samples when the scene's
// last render was not theirs.
struct RenderContext {
    nr_scene* scene = nullptr;
    nr_camera cam{};
    nr_params params{};
    uint32_t n_mats = 0;
    std::vector<float> theta;
};

struct RenderFunction {
    static std::vector<float> forward(RenderContext& ctx, Renderer& r, const Camera& cam, const Config& cfg,
                                      const std::vector<float>& theta) {
        if (!r.handle()) throw Error("forward before prepare");
        if (r.name() == "nrc") throw Error("the radiance-cache renderer has no adjoint (use ptnee or pt)");
        ctx.scene = r.handle();
        ctx.cam = cam.c();
        ctx.params = Renderer::params_of(cfg);
        ctx.params.integrator = r.name() == "pt" ? 1 : 0;
        ctx.params.batch_spp = 0;
        ctx.params.accumulate = 0;
        ctx.n_mats = uint32_t(theta.size() / 4);
        ctx.theta = theta;
        check(nr_scene_set_params(ctx.scene, theta.data(), ctx.n_mats));
        std::vector<float> img(size_t(ctx.params.width) * ctx.params.height * 3);
        check(nr_render_ptnee(ctx.scene, &ctx.cam, &ctx.params, img.data(), nullptr, nullptr, nullptr, nullptr, nullptr,
                              nullptr));
        return img;
    }
    static std::vector<double> backward(const RenderContext& ctx, const std::vector<float>& grad_image) {
        if (grad_image.size() != size_t(ctx.params.width) * ctx.params.height * 3)
            throw Error("grad_image must have W*H*3 entries");
        std::vector<double> g(size_t(ctx.n_mats) * 4);
        check(nr_scene_set_params(ctx.scene, ctx.theta.data(), ctx.n_mats));
        check(nr_adjoint_ptnee(ctx.scene, &ctx.cam, &ctx.params, grad_image.data(), g.data()));
        return g;
    }
};

}  // namespace nuray

#endif
