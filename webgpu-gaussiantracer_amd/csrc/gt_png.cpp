/*
 * gt_png.cpp - PNG file of an 8-bit RGBA image: what the reference's Save
 * button downloads (src/main.ts:606-611, canvas.toDataURL('image/png')).
 *
 * Written from the format specifications: PNG (RFC 2083: signature, IHDR
 * colour type 6 / depth 8, IDAT, IEND, CRC-32 of type + data), zlib
 * (RFC 1950: CMF/FLG header, Adler-32 trailer) and deflate (RFC 1951:
 * stored blocks of at most 65535 bytes).  Every scanline uses filter 0, so
 * the file holds the pixels verbatim.
 */
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/gt.h"
#include "gt_internal.h"

namespace {

struct Crc32 {
  uint32_t table[256];
  Crc32() {
    for (uint32_t i = 0; i < 256; ++i) {
      uint32_t c = i;
      for (int k = 0; k < 8; ++k) c = (c & 1u) ? 0xEDB88320u ^ (c >> 1) : (c >> 1);
      table[i] = c;
    }
  }
  uint32_t update(uint32_t crc, const uint8_t* p, size_t n) const {
    uint32_t c = ~crc;
    for (size_t i = 0; i < n; ++i) c = table[(c ^ p[i]) & 0xffu] ^ (c >> 8);
    return ~c;
  }
};

void put_be32(std::vector<uint8_t>& v, uint32_t x) {
  v.push_back((uint8_t)(x >> 24)); v.push_back((uint8_t)(x >> 16));
  v.push_back((uint8_t)(x >> 8)); v.push_back((uint8_t)x);
}

void put_chunk(std::vector<uint8_t>& file, const Crc32& crc, const char type[4], const std::vector<uint8_t>& data) {
  put_be32(file, (uint32_t)data.size());
  const size_t at = file.size();
  file.insert(file.end(), type, type + 4);
  file.insert(file.end(), data.begin(), data.end());
  put_be32(file, crc.update(0u, file.data() + at, 4 + data.size()));
}

}  // namespace

extern "C" int gt_write_png(const char* path, const uint8_t* rgba8, uint32_t width, uint32_t height) {
  if (!path) { gt_set_error("gt_write_png: path is NULL"); return GT_E_INVALID; }
  if (width == 0 || height == 0 || width > 16384 || height > 16384) {
    gt_set_error("gt_write_png: %u x %u outside 1..16384", width, height);
    return GT_E_INVALID;
  }
  if (!rgba8) { gt_set_error("gt_write_png: image is NULL"); return GT_E_INVALID; }
  static const Crc32 crc;
  const size_t row = (size_t)width * 4;
  const size_t raw_len = (row + 1) * height; /* filter byte + pixels per scanline */
  std::vector<uint8_t> z;
  z.reserve(raw_len + raw_len / 65535 * 5 + 16);
  z.push_back(0x78); /* CMF: deflate, 32K window */
  z.push_back(0x01); /* FLG: (0x78 * 256 + 0x01) % 31 == 0, no dictionary */
  uint32_t a = 1, b = 0; /* Adler-32 of the uncompressed stream */
  size_t done = 0, block_left = 0;
  auto emit = [&](uint8_t byte) {
    if (block_left == 0) {
      const size_t len = std::min<size_t>(65535, raw_len - done);
      z.push_back(done + len == raw_len ? 1u : 0u); /* BFINAL, BTYPE 00 (stored) */
      z.push_back((uint8_t)len); z.push_back((uint8_t)(len >> 8));
      z.push_back((uint8_t)~len); z.push_back((uint8_t)(~len >> 8));
      block_left = len;
    }
    z.push_back(byte);
    a = (a + byte) % 65521u;
    b = (b + a) % 65521u;
    --block_left;
    ++done;
  };
  for (uint32_t y = 0; y < height; ++y) {
    emit(0u); /* filter: none */
    const uint8_t* r = rgba8 + (size_t)y * row;
    for (size_t i = 0; i < row; ++i) emit(r[i]);
  }
  put_be32(z, (b << 16) | a);

  std::vector<uint8_t> file = {0x89, 'P', 'N', 'G', 0x0D, 0x0A, 0x1A, 0x0A};
  std::vector<uint8_t> ihdr;
  put_be32(ihdr, width);
  put_be32(ihdr, height);
  ihdr.insert(ihdr.end(), {8u, 6u, 0u, 0u, 0u}); /* depth 8, RGBA, deflate, filter 0, no interlace */
  put_chunk(file, crc, "IHDR", ihdr);
  put_chunk(file, crc, "IDAT", z);
  put_chunk(file, crc, "IEND", {});
  FILE* f = std::fopen(path, "wb");
  if (!f) { gt_set_error("gt_write_png: cannot open %s", path); return GT_E_IO; }
  const size_t wrote = std::fwrite(file.data(), 1, file.size(), f);
  const int closed = std::fclose(f);
  if (wrote != file.size() || closed != 0) { gt_set_error("gt_write_png: write to %s failed", path); return GT_E_IO; }
  return GT_OK;
}
