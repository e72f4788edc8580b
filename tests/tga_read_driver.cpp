// Host-only driver for the TGA reader test: reads one file with the
// library's tga_read and prints "ok" or the error message (exit status 0 / 1).
#include <cstdio>
#include <string>

#include "host_io.hpp"

int main(int argc, char **argv)
{
    if (argc != 2) return 2;
    fgl::HostTexture t;
    std::string err;
    if (!fgl::tga_read(argv[1], t, err)) {
        std::printf("%s\n", err.c_str());
        return 1;
    }
    std::printf("ok %d %d %d\n", t.w, t.h, t.bpp);
    return 0;
}
