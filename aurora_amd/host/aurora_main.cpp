// aurora_main.cpp -- the Aurora command line over the MI355X renderer.
//
// Mirrors Aurora/main.cpp:36-52: one scene file argument (plus, here, an
// optional config.json for the scene's $-variables), a usage message when the
// count is wrong, an error when the file does not open, else
// Renderer(file).render().  The reference logs those errors and still exits
// 0; this program exits 1 so scripts see them, and 2 when the library
// reports a failure (with its message).
#include <cstdio>
#include <fstream>

#include "aurora_gpu.hpp"

int main(int argc, char *argv[]) {
    if (argc != 2 && argc != 3) {
        std::fprintf(stderr, "usage: %s <filename> [config.json].\n", argv[0]);
        return 1;
    }
    {
        std::ifstream the_file(argv[1]);
        if (!the_file.is_open()) {
            std::fprintf(stderr, "Could not open file %s.\n", argv[1]);
            return 1;
        }
    }
    try {
        aurora_gpu::Renderer ren(argv[1], argc == 3 ? argv[2] : "");
        ren.render();
    } catch (const aurora_gpu::Error &e) {
        std::fprintf(stderr, "%s (code %d)\n", e.what(), e.code);
        return 2;
    }
    return 0;
}
