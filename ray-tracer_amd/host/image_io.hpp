// image_io.hpp — 8-bit RGB images and their encoders (binary PPM; PNG with
// zlib DEFLATE), the output formats of the `rt` CLI (src/io/image_io.cpp).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace rtm {

struct Image8 {  // row-major, top row first, RGB interleaved
  int width = 0, height = 0;
  std::vector<uint8_t> rgb;
};

std::vector<uint8_t> encode_png(const Image8& img);
void write_ppm(const std::string& path, const Image8& img);
void write_png(const std::string& path, const Image8& img);
// .png or .ppm by extension; anything else throws.
void write_image(const std::string& path, const Image8& img);

}  // namespace rtm
