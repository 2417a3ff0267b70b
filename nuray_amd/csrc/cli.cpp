// nuray_cli: the reference's command line (cmd.cpp:9-113) on the GPU path.
//
//   nuray_cli <scene description> --output <file.bmp> --params k=v k=v ...
//   e.g. nuray_cli tests/golden/cornell/CornellBox-Original.obj -s 100 --output result.bmp
//        --params renderer=ptnee imgw=320 imgh=240 spp=16 campos=0,120,200 cameuler=0,0,0 camfov=64
//
// Tokens before --output form the scene description line (Loader::fromSceneDescription); the
// params are Config items.  Camera(camfov, camasp, campos, cameuler) as cmd.cpp:98-100, envmap=<BMP
// or PPM> with envmap_gain.  renderer = ptnee (default), pt or nrc.
#include <iostream>
#include <string>
#include <vector>
#include "../../include/nuray.hpp"

int main(int argc, char** argv) {
    std::vector<std::string> args(argv + 1, argv + argc);
    nuray::Config config;
    std::string desc, output;
    int stage = 0;
    for (const std::string& s : args) {
        if (s == "--output") { stage = 1; continue; }
        if (s == "--params") { stage = 2; continue; }
        if (stage == 0) desc += s + " ";
        else if (stage == 1) output = s;
        else config.addItem(s);
    }
    if (desc.empty() || output.empty()) {
        std::cerr << "usage: nuray_cli <scene description> --output <file.bmp> --params k=v ..." << std::endl;
        return 2;
    }
    try {
        std::cout << "config list" << std::endl;
        config.print();
        nuray::Loader loader;
        loader.fromSceneDescription(desc);
        std::string name = config.getValueStr("renderer", "ptnee");
        if (name.empty()) name = "ptnee";
        nuray::Renderer r(name);
        const std::string env = config.getValueStr("envmap", "");
        if (!env.empty()) r.setEnvmap(nuray::Envmap(nuray::Texture::load(env), config.getValueFloat("envmap_gain", 1.0f)));
        r.prepare(loader);
        const nuray::vec3 e = config.getValueVec3("cameuler", nuray::vec3(0.0f));
        nuray::Camera cam(config.getValueFloat("camfov", 90.0f), config.getValueFloat("camasp", 1.0f),
                          config.getValueVec3("campos", nuray::vec3(0.0f)), e[0], e[1], e[2]);
        nuray::Image img;
        r.render(cam, img, config, [](float x) { std::cout << "Process: " << 100.0f * x << "% \t"; });
        std::cout << std::endl;
        img.saveBMP(output);
        std::cout << "finish:) " << loader.getTriangles().size() << " triangles -> " << output << std::endl;
    } catch (const std::exception& ex) {
        std::cerr << "nuray_cli: " << ex.what() << std::endl;
        return 1;
    }
    return 0;
}
