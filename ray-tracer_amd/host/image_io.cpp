// image_io.cpp — PPM (P6) and PNG writers. PNG: IHDR (8-bit RGB), one IDAT
// holding the zlib stream of the filter-0 scanlines, IEND; CRC-32 and
// DEFLATE come from the system zlib.
#include "image_io.hpp"

#include <zlib.h>

#include <fstream>
#include <stdexcept>

namespace rtm {

namespace {

void check(const Image8& img) {
  if (img.width <= 0 || img.height <= 0 || img.rgb.size() != (size_t)img.width * (size_t)img.height * 3)
    throw std::runtime_error("invalid image dimensions");
}

void put_be32(std::vector<uint8_t>& out, uint32_t v) {
  out.push_back((uint8_t)(v >> 24));
  out.push_back((uint8_t)(v >> 16));
  out.push_back((uint8_t)(v >> 8));
  out.push_back((uint8_t)v);
}

void chunk(std::vector<uint8_t>& out, const char type[4], const uint8_t* data, size_t n) {
  put_be32(out, (uint32_t)n);
  const size_t start = out.size();
  out.insert(out.end(), type, type + 4);
  if (n) out.insert(out.end(), data, data + n);
  const uint32_t crc = (uint32_t)crc32(0L, out.data() + start, (uInt)(n + 4));
  put_be32(out, crc);
}

bool ends_with(const std::string& s, const std::string& suf) {
  return s.size() >= suf.size() && s.compare(s.size() - suf.size(), suf.size(), suf) == 0;
}

}  // namespace

std::vector<uint8_t> encode_png(const Image8& img) {
  check(img);
  const size_t row = (size_t)img.width * 3;
  std::vector<uint8_t> raw((row + 1) * (size_t)img.height);
  for (int y = 0; y < img.height; ++y) {
    raw[(size_t)y * (row + 1)] = 0;  // filter: none
    std::copy(img.rgb.begin() + (ptrdiff_t)((size_t)y * row), img.rgb.begin() + (ptrdiff_t)((size_t)(y + 1) * row),
              raw.begin() + (ptrdiff_t)((size_t)y * (row + 1) + 1));
  }
  uLongf zlen = compressBound((uLong)raw.size());
  std::vector<uint8_t> z(zlen);
  if (compress2(z.data(), &zlen, raw.data(), (uLong)raw.size(), 6) != Z_OK)
    throw std::runtime_error("PNG encoding failed");
  z.resize(zlen);
  std::vector<uint8_t> out = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
  uint8_t ihdr[13];
  const uint32_t w = (uint32_t)img.width, h = (uint32_t)img.height;
  const uint8_t hdr[13] = {(uint8_t)(w >> 24), (uint8_t)(w >> 16), (uint8_t)(w >> 8), (uint8_t)w,
                           (uint8_t)(h >> 24), (uint8_t)(h >> 16), (uint8_t)(h >> 8), (uint8_t)h,
                           8, 2, 0, 0, 0};  // depth 8, colour type 2 (RGB)
  std::copy(hdr, hdr + 13, ihdr);
  chunk(out, "IHDR", ihdr, 13);
  chunk(out, "IDAT", z.data(), z.size());
  chunk(out, "IEND", nullptr, 0);
  return out;
}

void write_ppm(const std::string& path, const Image8& img) {
  check(img);
  std::ofstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot write " + path);
  f << "P6\n" << img.width << ' ' << img.height << "\n255\n";
  f.write(reinterpret_cast<const char*>(img.rgb.data()), (std::streamsize)img.rgb.size());
  if (!f) throw std::runtime_error("write failed: " + path);
}

void write_png(const std::string& path, const Image8& img) {
  const auto bytes = encode_png(img);
  std::ofstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot write " + path);
  f.write(reinterpret_cast<const char*>(bytes.data()), (std::streamsize)bytes.size());
  if (!f) throw std::runtime_error("write failed: " + path);
}

void write_image(const std::string& path, const Image8& img) {
  if (ends_with(path, ".ppm")) write_ppm(path, img);
  else if (ends_with(path, ".png")) write_png(path, img);
  else throw std::runtime_error("unsupported output extension (use .png or .ppm): " + path);
}

}  // namespace rtm
