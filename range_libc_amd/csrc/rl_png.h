// PNG reading and writing for occupancy maps (host code, zlib for the
// compressed stream).
//
// OMap(filename, threshold) (RangeLib.h:158-201) decodes a PNG to 8-bit RGBA
// and marks a cell occupied where the gray value of its pixel is below the
// threshold; OMap::save (RangeLib.h:264-291) writes an opaque RGBA image, black
// where occupied.  This decoder produces the same 8-bit RGBA samples a PNG
// decoder must for every standard format:
//   colour types 0 (gray), 2 (RGB), 3 (palette), 4 (gray + alpha), 6 (RGBA);
//   bit depths 1, 2, 4, 8, 16 where the format allows them; 16-bit samples
//   keep their most significant byte; gray samples below 8 bits are scaled to
//   0..255 (v * 255 / (2^depth - 1)); tRNS gives alpha (only the colour
//   channels enter the gray value);
//   scanline filters None, Sub, Up, Average, Paeth; Adam7 interlacing;
//   chunk CRCs and the zlib checksum are verified.
#pragma once
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <zlib.h>

#include <string>
#include <vector>

namespace rlpng {

struct Image {
  unsigned width = 0, height = 0;
  std::vector<uint8_t> rgba;  // 4 * width * height, row-major
};

inline uint32_t be32(const uint8_t* p) {
  return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | (uint32_t)p[3];
}

inline void put_be32(std::vector<uint8_t>& v, uint32_t x) {
  v.push_back((uint8_t)(x >> 24));
  v.push_back((uint8_t)(x >> 16));
  v.push_back((uint8_t)(x >> 8));
  v.push_back((uint8_t)x);
}

inline bool read_file(const char* path, std::vector<uint8_t>* out) {
  FILE* f = fopen(path, "rb");
  if (!f) return false;
  uint8_t buf[65536];
  size_t n;
  while ((n = fread(buf, 1, sizeof(buf), f)) > 0) out->insert(out->end(), buf, buf + n);
  bool ok = !ferror(f);
  fclose(f);
  return ok;
}

inline int paeth(int a, int b, int c) {
  const int p = a + b - c;
  const int pa = p > a ? p - a : a - p, pb = p > b ? p - b : b - p, pc = p > c ? p - c : c - p;
  if (pa <= pb && pa <= pc) return a;
  return pb <= pc ? b : c;
}

// Reverses the filters of `rows` scanlines of `stride` bytes (each preceded by
// its filter-type byte) in place; bpp = bytes per complete pixel, at least 1.
inline bool unfilter(uint8_t* data, size_t rows, size_t stride, size_t bpp, std::string* err) {
  std::vector<uint8_t> prev(stride, 0);
  for (size_t r = 0; r < rows; ++r) {
    uint8_t* line = data + r * (stride + 1);
    const int type = line[0];
    uint8_t* cur = line + 1;
    for (size_t i = 0; i < stride; ++i) {
      const int a = i >= bpp ? cur[i - bpp] : 0;
      const int b = prev[i];
      const int c = i >= bpp ? prev[i - bpp] : 0;
      int pred;
      switch (type) {
        case 0: pred = 0; break;
        case 1: pred = a; break;
        case 2: pred = b; break;
        case 3: pred = (a + b) >> 1; break;
        case 4: pred = paeth(a, b, c); break;
        default:
          *err = "unknown scanline filter type " + std::to_string(type);
          return false;
      }
      cur[i] = (uint8_t)(cur[i] + pred);
    }
    memcpy(prev.data(), cur, stride);
  }
  return true;
}

struct Format {
  int depth, ctype, channels;
  std::vector<uint8_t> palette;  // RGB triples
  std::vector<uint8_t> trns;     // raw tRNS chunk
};

// Sample `s` (0-based channel index within the pixel) of pixel x of an
// unfiltered scanline, as a value of `depth` bits.
inline unsigned sample(const uint8_t* line, unsigned x, int s, const Format& f) {
  const unsigned idx = x * (unsigned)f.channels + (unsigned)s;
  if (f.depth == 8) return line[idx];
  if (f.depth == 16) return ((unsigned)line[2 * idx] << 8) | line[2 * idx + 1];
  const unsigned bit = idx * (unsigned)f.depth;
  const unsigned byte = line[bit >> 3];
  const unsigned shift = 8u - (unsigned)f.depth - (bit & 7u);
  return (byte >> shift) & ((1u << f.depth) - 1u);
}

inline uint8_t to8(unsigned v, int depth) {
  if (depth == 16) return (uint8_t)(v >> 8);
  if (depth == 8) return (uint8_t)v;
  return (uint8_t)(v * 255u / ((1u << depth) - 1u));
}

inline void convert_pixel(const uint8_t* line, unsigned x, const Format& f, uint8_t* o) {
  const bool has_key = !f.trns.empty();
  switch (f.ctype) {
    case 0: {
      const unsigned g = sample(line, x, 0, f);
      o[0] = o[1] = o[2] = to8(g, f.depth);
      o[3] = (has_key && f.trns.size() >= 2 && g == ((unsigned)f.trns[0] << 8 | f.trns[1])) ? 0 : 255;
      break;
    }
    case 2: {
      unsigned v[3];
      for (int s = 0; s < 3; ++s) v[s] = sample(line, x, s, f);
      for (int s = 0; s < 3; ++s) o[s] = to8(v[s], f.depth);
      bool key = has_key && f.trns.size() >= 6;
      for (int s = 0; s < 3 && key; ++s) key = v[s] == ((unsigned)f.trns[2 * s] << 8 | f.trns[2 * s + 1]);
      o[3] = key ? 0 : 255;
      break;
    }
    case 3: {
      const unsigned i = sample(line, x, 0, f);
      if (3 * i + 2 < f.palette.size()) {
        o[0] = f.palette[3 * i];
        o[1] = f.palette[3 * i + 1];
        o[2] = f.palette[3 * i + 2];
      } else {
        o[0] = o[1] = o[2] = 0;
      }
      o[3] = i < f.trns.size() ? f.trns[i] : 255;
      break;
    }
    case 4:
      o[0] = o[1] = o[2] = to8(sample(line, x, 0, f), f.depth);
      o[3] = to8(sample(line, x, 1, f), f.depth);
      break;
    default:  // 6
      for (int s = 0; s < 4; ++s) o[s] = to8(sample(line, x, s, f), f.depth);
      break;
  }
}

// Decodes a PNG file to 8-bit RGBA.  Returns false with a message on failure.
inline bool decode_file(const char* path, Image* img, std::string* err) {
  std::vector<uint8_t> file;
  if (!read_file(path, &file)) {
    *err = std::string("cannot read ") + path;
    return false;
  }
  static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
  if (file.size() < 8 || memcmp(file.data(), sig, 8) != 0) {
    *err = "not a PNG file (bad signature)";
    return false;
  }
  Format f{0, 0, 0, {}, {}};
  unsigned W = 0, H = 0;
  int interlace = 0;
  bool have_ihdr = false, have_iend = false;
  std::vector<uint8_t> idat;
  size_t pos = 8;
  while (pos + 12 <= file.size() && !have_iend) {
    const uint32_t len = be32(&file[pos]);
    if (len > 0x7fffffffu || pos + 12 + (size_t)len > file.size()) {
      *err = "chunk extends past the end of the file";
      return false;
    }
    const uint8_t* type = &file[pos + 4];
    const uint8_t* d = &file[pos + 8];
    const uint32_t crc = be32(&file[pos + 8 + len]);
    if ((uint32_t)crc32(crc32(0L, Z_NULL, 0), type, 4 + len) != crc) {
      *err = std::string("CRC mismatch in chunk ") + std::string((const char*)type, 4);
      return false;
    }
    if (!memcmp(type, "IHDR", 4)) {
      if (len != 13) {
        *err = "bad IHDR";
        return false;
      }
      W = be32(d);
      H = be32(d + 4);
      f.depth = d[8];
      f.ctype = d[9];
      interlace = d[12];
      if (d[10] != 0 || d[11] != 0 || interlace > 1) {
        *err = "unsupported compression, filter or interlace method";
        return false;
      }
      have_ihdr = true;
    } else if (!memcmp(type, "PLTE", 4)) {
      f.palette.assign(d, d + len);
    } else if (!memcmp(type, "tRNS", 4)) {
      f.trns.assign(d, d + len);
    } else if (!memcmp(type, "IDAT", 4)) {
      idat.insert(idat.end(), d, d + len);
    } else if (!memcmp(type, "IEND", 4)) {
      have_iend = true;
    } else if (!(type[0] & 0x20)) {
      *err = std::string("unknown critical chunk ") + std::string((const char*)type, 4);
      return false;
    }
    pos += 12 + (size_t)len;
  }
  if (!have_ihdr || !have_iend || idat.empty()) {
    *err = "missing IHDR, IDAT or IEND";
    return false;
  }
  if (W == 0 || H == 0 || W > (1u << 24) || H > (1u << 24)) {
    *err = "image size out of range";
    return false;
  }
  const int d = f.depth;
  switch (f.ctype) {
    case 0: f.channels = 1; if (d != 1 && d != 2 && d != 4 && d != 8 && d != 16) f.channels = 0; break;
    case 2: f.channels = 3; if (d != 8 && d != 16) f.channels = 0; break;
    case 3: f.channels = 1; if (d != 1 && d != 2 && d != 4 && d != 8) f.channels = 0; break;
    case 4: f.channels = 2; if (d != 8 && d != 16) f.channels = 0; break;
    case 6: f.channels = 4; if (d != 8 && d != 16) f.channels = 0; break;
    default: f.channels = 0;
  }
  if (!f.channels) {
    *err = "invalid colour type / bit depth combination";
    return false;
  }
  if (f.ctype == 3 && f.palette.empty()) {
    *err = "palette image without PLTE";
    return false;
  }
  const size_t bits_pp = (size_t)f.channels * (size_t)d;
  const size_t bpp = bits_pp < 8 ? 1 : bits_pp / 8;
  // pass geometry: Adam7 (7 passes) or one pass covering the image
  static const unsigned ax0[7] = {0, 4, 0, 2, 0, 1, 0}, ay0[7] = {0, 0, 4, 0, 2, 0, 1};
  static const unsigned adx[7] = {8, 8, 4, 4, 2, 2, 1}, ady[7] = {8, 8, 8, 4, 4, 2, 2};
  const int passes = interlace ? 7 : 1;
  size_t expect = 0;
  unsigned pw[7], ph[7];
  for (int p = 0; p < passes; ++p) {
    pw[p] = interlace ? (W > ax0[p] ? (W - ax0[p] + adx[p] - 1) / adx[p] : 0) : W;
    ph[p] = interlace ? (H > ay0[p] ? (H - ay0[p] + ady[p] - 1) / ady[p] : 0) : H;
    if (pw[p] && ph[p]) expect += (size_t)ph[p] * (1 + (pw[p] * bits_pp + 7) / 8);
  }
  std::vector<uint8_t> raw(expect);
  uLongf got = (uLongf)expect;
  const int zrc = uncompress(raw.data(), &got, idat.data(), (uLong)idat.size());
  if (zrc != Z_OK || got != expect) {
    *err = "zlib stream: " + std::string(zrc == Z_OK ? "wrong decompressed size" : zError(zrc));
    return false;
  }
  img->width = W;
  img->height = H;
  img->rgba.assign((size_t)W * H * 4, 0);
  size_t off = 0;
  for (int p = 0; p < passes; ++p) {
    if (!pw[p] || !ph[p]) continue;
    const size_t stride = (pw[p] * bits_pp + 7) / 8;
    if (!unfilter(&raw[off], ph[p], stride, bpp, err)) return false;
    for (unsigned y = 0; y < ph[p]; ++y) {
      const uint8_t* line = &raw[off + y * (stride + 1) + 1];
      for (unsigned x = 0; x < pw[p]; ++x) {
        const unsigned X = interlace ? ax0[p] + x * adx[p] : x;
        const unsigned Y = interlace ? ay0[p] + y * ady[p] : y;
        convert_pixel(line, x, f, &img->rgba[4 * ((size_t)Y * W + X)]);
      }
    }
    off += (size_t)ph[p] * (stride + 1);
  }
  return true;
}

inline void chunk(std::vector<uint8_t>& out, const char* type, const std::vector<uint8_t>& data) {
  put_be32(out, (uint32_t)data.size());
  const size_t start = out.size();
  out.insert(out.end(), type, type + 4);
  out.insert(out.end(), data.begin(), data.end());
  put_be32(out, (uint32_t)crc32(crc32(0L, Z_NULL, 0), &out[start], (uInt)(4 + data.size())));
}

// Writes 8-bit RGBA rows (4 * width * height bytes) as a PNG file.
inline bool encode_file(const char* path, const uint8_t* rgba, unsigned W, unsigned H, std::string* err) {
  std::vector<uint8_t> rows;
  rows.reserve((size_t)H * (1 + 4 * (size_t)W));
  for (unsigned y = 0; y < H; ++y) {
    rows.push_back(0);  // filter: none
    rows.insert(rows.end(), rgba + 4 * (size_t)y * W, rgba + 4 * ((size_t)y + 1) * W);
  }
  uLongf zlen = compressBound((uLong)rows.size());
  std::vector<uint8_t> z(zlen);
  if (compress2(z.data(), &zlen, rows.data(), (uLong)rows.size(), 6) != Z_OK) {
    *err = "zlib compression failed";
    return false;
  }
  z.resize(zlen);
  std::vector<uint8_t> out = {137, 80, 78, 71, 13, 10, 26, 10};
  std::vector<uint8_t> ihdr;
  put_be32(ihdr, W);
  put_be32(ihdr, H);
  ihdr.insert(ihdr.end(), {8, 6, 0, 0, 0});
  chunk(out, "IHDR", ihdr);
  chunk(out, "IDAT", z);
  chunk(out, "IEND", {});
  FILE* fp = fopen(path, "wb");
  if (!fp) {
    *err = std::string("cannot write ") + path;
    return false;
  }
  const bool ok = fwrite(out.data(), 1, out.size(), fp) == out.size();
  if (fclose(fp) != 0 || !ok) {
    *err = std::string("write failed: ") + path;
    return false;
  }
  return true;
}

}  // namespace rlpng
