// host_api_check.cpp -- drives the C++ host classes of aurora_gpu.hpp from a
// C++ program, for tests/test_gpu_host_cpp.py.
//
//   host_api_check <scene> <config.json> <rays.bin> <out.bin>
//
// rays.bin holds n aur_ray records.  The program loads the scene with
// Renderer::parseSceneDescription, checks that a query before
// buildRenderEnvironment throws aurora_gpu::Error with AUR_ESTATE, builds,
// then writes n aur_hit records of AccelerationStructure::intersect followed
// by n bytes of intersectBinary to out.bin.
#include <cstdio>
#include <fstream>
#include <iterator>
#include <vector>

#include "../../aurora_amd/host/aurora_gpu.hpp"

int main(int argc, char *argv[]) {
    if (argc != 5) {
        std::fprintf(stderr, "usage: %s <scene> <config.json> <rays.bin> <out.bin>\n", argv[0]);
        return 1;
    }
    std::ifstream in(argv[3], std::ios::binary);
    std::vector<char> bytes((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    if (!in.good() && !in.eof()) return 1;
    std::vector<aur_ray> rays(bytes.size() / sizeof(aur_ray));
    std::copy(bytes.begin(), bytes.begin() + rays.size() * sizeof(aur_ray), reinterpret_cast<char *>(rays.data()));
    try {
        aurora_gpu::Renderer ren(argv[1], argv[2]);
        ren.parseSceneDescription();
        aurora_gpu::AccelerationStructure accel(ren.context());
        try {
            accel.intersect(rays);
            std::fprintf(stderr, "a query before the build did not throw\n");
            return 3;
        } catch (const aurora_gpu::Error &e) {
            if (e.code != AUR_ESTATE) {
                std::fprintf(stderr, "query before the build: code %d (%s)\n", e.code, e.what());
                return 3;
            }
            std::printf("state error: %s\n", e.what());
        }
        ren.buildRenderEnvironment();
        const std::vector<aur_hit> hits = accel.intersect(rays);
        const std::vector<bool> blocked = accel.intersectBinary(rays);
        std::ofstream out(argv[4], std::ios::binary);
        out.write(reinterpret_cast<const char *>(hits.data()), (std::streamsize)(hits.size() * sizeof(aur_hit)));
        for (bool b : blocked) out.put(b ? 1 : 0);
        if (!out) return 1;
        std::printf("traced %zu rays\n", rays.size());
    } catch (const aurora_gpu::Error &e) {
        std::fprintf(stderr, "%s (code %d)\n", e.what(), e.code);
        return 2;
    }
    return 0;
}
