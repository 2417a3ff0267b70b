// Host-only checks of the C++ facade (include/nuray.hpp): Config semantics, Loader output dumped
// for comparison with an independent parse, BMP write / read round trip.  No GPU calls.
#include <cstdio>
#include <fstream>
#include "../../include/nuray.hpp"

static int fails = 0;
#define EXPECT(c) do { if (!(c)) { std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #c); ++fails; } } while (0)

int main(int argc, char** argv) {
    if (argc < 4) return 2;
    const std::string obj = argv[1], dump = argv[2], tmpbmp = argv[3];
    // Config (config.cpp:12-104)
    nuray::Config c("imgw=320 spp=abc campos=0,120,200 prr=0.5 renderer=ptnee bad");
    EXPECT(c.getValueInt("imgw", 1) == 320);
    EXPECT(c.getValueInt("spp", 7) == 7);               // unparsable -> default
    EXPECT(c.getValueInt("missing", 9) == 9);
    EXPECT(c.getValueFloat("prr", 0.8f) == 0.5f);
    EXPECT(c.getValueStr("missing", "ptnee").empty());   // "" for a missing key, whatever the default
    EXPECT(c.getValueStr("renderer", "") == "ptnee");
    nuray::vec3 p = c.getValueVec3("campos", nuray::vec3(0.0f));
    EXPECT(p[0] == 0.0f && p[1] == 120.0f && p[2] == 200.0f);
    nuray::vec3 d = c.getValueVec3("prr", nuray::vec3(1.0f, 2.0f, 3.0f));   // one number: default
    EXPECT(d[0] == 1.0f && d[2] == 3.0f);
    c.addItem("spp=16");
    EXPECT(c.getValueInt("spp", 7) == 16);
    // Loader: scene description line with scale and position
    nuray::Loader l;
    l.fromSceneDescription(obj + " -s 100 -p 1 2 3");
    const auto& tris = l.getTriangles();
    const auto& mats = l.materials();
    std::ofstream os(dump, std::ios::binary);
    const int32_t nt = int32_t(tris.size()), nm = int32_t(mats.size());
    os.write(reinterpret_cast<const char*>(&nt), 4);
    os.write(reinterpret_cast<const char*>(&nm), 4);
    for (const auto& t : tris) {
        for (int k = 0; k < 3; ++k) os.write(reinterpret_cast<const char*>(t.p[k].v), 12);
        for (int k = 0; k < 3; ++k) os.write(reinterpret_cast<const char*>(t.n[k].v), 12);
        os.write(reinterpret_cast<const char*>(&t.mat), 4);
    }
    for (const auto& m : mats) {
        const float row[8] = {float(m.type), m.kd[0], m.kd[1], m.kd[2], m.alpha, m.ior, float(m.tex), 0.0f};
        os.write(reinterpret_cast<const char*>(row), 32);
    }
    os.close();
    // forcing material
    nuray::Loader l2;
    l2.fromSceneDescription(obj + " -m=ggxr 0.5 0.6 0.7 0.3");
    EXPECT(l2.materials().size() == 1 && l2.materials()[0].type == NR_MAT_GGXREFL && l2.materials()[0].alpha == 0.3f);
    for (const auto& t : l2.getTriangles()) EXPECT(t.mat == 0);
    // BMP round trip (odd width: row padding)
    nuray::Image img;
    img.resize(5, 3);
    for (size_t i = 0; i < img.rgb.size(); ++i) img.rgb[i] = uint8_t(i * 7 + 3);
    img.saveBMP(tmpbmp);
    nuray::Texture t = nuray::Texture::load(tmpbmp);
    EXPECT(t.w == 5 && t.h == 3);
    bool same = true;
    for (int i = 0; i < 15; ++i)
        for (int k = 0; k < 3; ++k) same = same && t.rgba[4 * i + k] == img.rgb[3 * i + k];
    EXPECT(same);
    // forcing glass; unsupported things fail loudly
    nuray::Loader l3;
    l3.fromSceneDescription(obj + " -m=glass 0.9 0.9 0.9 1.5");
    EXPECT(l3.materials().size() == 1 && l3.materials()[0].type == NR_MAT_GLASS && l3.materials()[0].ior == 1.5f);
    // Ks > 0 in the MTL: Blinn-Phong with Kd, Ks and the exponent Ns (loader.cpp:126-140)
    if (argc > 4) {
        nuray::Loader l5;
        l5.fromSceneDescription(std::string(argv[4]) + " -s 100");
        int bp = 0;
        for (const auto& m : l5.materials())
            if (m.type == NR_MAT_BLINNPHONG) {
                ++bp;
                EXPECT(m.kd[0] == 0.01f && m.ks[0] == 0.95f && m.ks[2] == 0.95f && m.alpha == 100000.0f && m.tex_ks < 0);
            }
        EXPECT(bp == 1);
    }
    bool threw = false;
    try { nuray::Loader l4; l4.fromSceneDescription(obj + "_missing.obj"); } catch (const nuray::Error&) { threw = true; }
    EXPECT(threw);
    std::printf("%s %d triangles %d materials\n", fails ? "FAILED" : "OK", nt, nm);
    return fails ? 1 : 0;
}
