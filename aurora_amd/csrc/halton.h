// halton.h -- the Halton QMC sample source (the reference's optional
// Aurora/external/halton sampler: Faure-permuted Halton points, enumerated
// per pixel with Halton_enum; camera.cpp:39-45 under USE_HALTON).
//
// A path's sample slots come from one Halton point: its index is the i-th
// point (i = the pass) whose first two coordinates fall in the path's pixel;
// slots 0 and 1 are the camera's offsets inside the pixel, slot k in [2, 256)
// is dimension k.  Slots past 255 fall back to the counter-based hash.
#pragma once

#include <stdint.h>

#include <vector>

namespace aur {

#define AUR_HALTON_DIMS 256

// One dimension (base b): blocks of g digits (b^g <= 500) through a permuted
// lookup table; n blocks cover (b^g)^n < 2^32.
struct alignas(16) DHaltonDim {
    uint32_t block;    // b^g
    uint32_t n_blocks; // n
    uint32_t top;      // (b^g)^(n-1)
    uint32_t off;      // first entry in the table pool
    float scale;       // float((1 - 2^-22) / (b^g)^n)
    uint32_t pad[3];
};

// Per-scene view: the dimension records and tables, plus the per-pixel
// enumeration of the image's resolution.
struct alignas(16) HaltonView {
    const DHaltonDim *dims;
    const uint16_t *tables;
    int width;
    uint32_t mx, my, inc, p2, p3;
    float sx, sy;  // 2^p2, 3^p3
};

// Host tables: dims[256] and the concatenated permutation tables.
void halton_tables(std::vector<DHaltonDim> *dims, std::vector<uint16_t> *tables);
// The enumeration constants of a width x height image (Halton_enum's constructor).
void halton_enum(int width, int height, HaltonView *v);

}  // namespace aur
