// exr_io.h -- a small OpenEXR (scanline, single part) reader and writer.
//
// Replaces the OpenEXR library calls of the reference: the display driver's
// `Imf::RgbaOutputFile` (src/frontend/openexrDisplay.cpp:132-169: FP16 RGBA,
// ZIP compression, string metadata, written to `<file>_tmp` and renamed) and
// the texture loader's `Imf::RgbaInputFile` (src/frontend/texture2D.cpp:24-46).
// Supported: channels R, G, B, A (any subset; a file with Y and no R, G, B
// reads as grey) as HALF, FLOAT or UINT, compression
// NONE, ZIPS (one scanline per chunk) and ZIP (16 scanlines per chunk), the
// INCREASING_Y line order.  Layout follows the published OpenEXR file format
// (magic 20000630, version 2, attribute list, chunk offset table, chunks).
#pragma once

#include <stdint.h>

#include <string>
#include <utility>
#include <vector>

namespace aur {

enum { EXR_NONE = 0, EXR_ZIPS = 2, EXR_ZIP = 3 };

uint16_t float_to_half(float f);
float half_to_float(uint16_t h);

// rgba: width * height * 4 floats, row 0 first.  half = store FP16 channels.
void exr_write(const std::string &path, int width, int height, const float *rgba, int compression, bool half,
               const std::vector<std::pair<std::string, std::string>> &meta);

struct ExrImage {
    int width = 0, height = 0;
    std::vector<float> rgba;  // missing channels: 0 for R, G, B and 1 for A
    std::vector<std::pair<std::string, std::string>> meta;  // string attributes
};
ExrImage exr_read(const std::string &path);

}  // namespace aur
