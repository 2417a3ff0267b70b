// Decodes an image with the C++ facade's texture loader (include/nuray.hpp, nuray::Texture::load)
// and writes {int32 w, int32 h, w*h*4 RGBA8 bytes} to the output file; prints the loader's error
// and exits 1 when it refuses the file.  Host only.
#include <cstdio>
#include <fstream>
#include "../../include/nuray.hpp"

int main(int argc, char** argv) {
    if (argc < 3) return 2;
    try {
        const nuray::Texture t = nuray::Texture::load(argv[1]);
        std::ofstream os(argv[2], std::ios::binary);
        const int32_t wh[2] = {t.w, t.h};
        os.write(reinterpret_cast<const char*>(wh), 8);
        os.write(reinterpret_cast<const char*>(t.rgba.data()), std::streamsize(t.rgba.size()));
        return os ? 0 : 1;
    } catch (const std::exception& e) {
        std::printf("%s\n", e.what());
        return 1;
    }
}
