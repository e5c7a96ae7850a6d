// tiff_reader.hpp — frame ingest for the tracking procedure: baseline TIFF files decoded
// on the host into the reference's gray levels.
//
// The reference reads a frame with libtiff's TIFFReadRGBAImage (8-bit R, G, B per pixel)
// and stores, for the pixel in row y from the top and column x,
//   gray = (uint16_t)(65535.0 * (0.299 r + 0.587 g + 0.114 b)),  r = R * (1.0 / 255.0), ...
// (src/p-tracker.cpp:86-201; the raster is bottom-up and the reference flips it back, so
// the net orientation is the file's top-down row order). 0.299 + 0.587 + 0.114 rounds to
// 0.9999999999999999 in double, so full white is 65534. The frame is returned row-major,
// element (x, y) at frame[y * width + x], as libptgpu takes it.
//
// Supported: strips (chunky or planar), compression none (1), LZW (5), Deflate (8, 32946)
// and PackBits (32773), horizontal predictor (2) with LZW and Deflate, both byte orders, photometric
// MinIsWhite / MinIsBlack / RGB / Palette, 8- and 16-bit samples, orientation top-left.
// The 8-bit values libtiff's RGBA interface gives are reproduced: 8-bit samples as they
// are, 16-bit gray as v >> 8, 16-bit RGB as (v + 128) / 257, 16-bit colormaps as c >> 8
// (measured against libtiff 4 as installed with this image; the reference's libtiff
// version is not pinned).
#pragma once

#include <zlib.h>

#include <cstdint>
#include <cstring>
#include <fstream>
#include <iterator>
#include <stdexcept>
#include <string>
#include <vector>

namespace ptgpu {

struct tiff_frame {
  int width = 0, height = 0;
  std::vector<uint16_t> gray;  // row-major, top row first
  std::string dateTime;        // TIFFTAG_DATETIME, or "dateTime unknown" as the reference
};

namespace tiff_detail {

struct Reader {
  const std::vector<uint8_t>& b;
  bool big;
  uint16_t u16(size_t o) const {
    if (o + 2 > b.size()) throw std::runtime_error("TIFF: truncated file");
    return big ? (uint16_t)(b[o] << 8 | b[o + 1]) : (uint16_t)(b[o + 1] << 8 | b[o]);
  }
  uint32_t u32(size_t o) const {
    if (o + 4 > b.size()) throw std::runtime_error("TIFF: truncated file");
    return big ? ((uint32_t)b[o] << 24 | (uint32_t)b[o + 1] << 16 | (uint32_t)b[o + 2] << 8 | b[o + 3])
               : ((uint32_t)b[o + 3] << 24 | (uint32_t)b[o + 2] << 16 | (uint32_t)b[o + 1] << 8 | b[o]);
  }
};

struct Entry {
  uint16_t type = 0;
  uint32_t count = 0;
  size_t data = 0;  // offset of the value bytes
};

inline size_t type_size(uint16_t t) {
  switch (t) {
    case 1: case 2: case 6: case 7: return 1;
    case 3: case 8: return 2;
    case 4: case 9: case 11: return 4;
    case 5: case 10: case 12: return 8;
    default: return 0;
  }
}

inline uint32_t value(const Reader& r, const Entry& e, uint32_t i) {
  if (i >= e.count) throw std::runtime_error("TIFF: tag value index out of range");
  if (e.type == 3) return r.u16(e.data + 2 * i);
  if (e.type == 4) return r.u32(e.data + 4 * i);
  if (e.type == 1) return r.b.at(e.data + i);
  throw std::runtime_error("TIFF: unexpected tag type");
}

// TIFF LZW (MSB-first codes, 9 to 12 bits, code width grows one code early).
inline void lzw_decode(const uint8_t* in, size_t n, std::vector<uint8_t>& out) {
  std::vector<uint16_t> prefix(4096);
  std::vector<uint8_t> suffix(4096), first(4096);
  std::vector<uint16_t> len(4096);
  for (int i = 0; i < 256; i++) {
    prefix[i] = 0xffff;
    suffix[i] = first[i] = (uint8_t)i;
    len[i] = 1;
  }
  size_t bitpos = 0;
  const size_t nbits = n * 8;
  int width = 9, next = 258, old = -1;
  std::vector<uint8_t> tmp(4096);
  auto emit = [&](int code) {
    int l = len[code];
    for (int c = code, k = l - 1; k >= 0; k--, c = prefix[c]) tmp[k] = suffix[c];
    out.insert(out.end(), tmp.begin(), tmp.begin() + l);
  };
  while (bitpos + width <= nbits) {
    int code = 0;
    for (int k = 0; k < width; k++, bitpos++) code = (code << 1) | ((in[bitpos >> 3] >> (7 - (bitpos & 7))) & 1);
    if (code == 257) break;
    if (code == 256) {
      width = 9;
      next = 258;
      old = -1;
      continue;
    }
    if (old < 0) {
      if (code > 255) throw std::runtime_error("TIFF: bad LZW stream");
      emit(code);
      old = code;
      continue;
    }
    if (code < next) {
      emit(code);
      if (next < 4096) {
        prefix[next] = (uint16_t)old;
        suffix[next] = first[code];
        first[next] = first[old];
        len[next] = (uint16_t)(len[old] + 1);
        next++;
      }
    } else if (code == next && next < 4096) {
      prefix[next] = (uint16_t)old;
      suffix[next] = first[old];
      first[next] = first[old];
      len[next] = (uint16_t)(len[old] + 1);
      next++;
      emit(code);
    } else {
      throw std::runtime_error("TIFF: bad LZW code");
    }
    old = code;
    if (next >= (1 << width) - 1 && width < 12) width++;
  }
}

inline void packbits_decode(const uint8_t* in, size_t n, std::vector<uint8_t>& out) {
  size_t i = 0;
  while (i < n) {
    const int c = (int8_t)in[i++];
    if (c >= 0) {
      const size_t k = (size_t)c + 1;
      if (i + k > n) throw std::runtime_error("TIFF: bad PackBits run");
      out.insert(out.end(), in + i, in + i + k);
      i += k;
    } else if (c != -128) {
      if (i >= n) throw std::runtime_error("TIFF: bad PackBits run");
      out.insert(out.end(), (size_t)(1 - c), in[i++]);
    }
  }
}

inline void deflate_decode(const uint8_t* in, size_t n, size_t expect, std::vector<uint8_t>& out) {
  const size_t o = out.size();
  out.resize(o + expect);
  uLongf got = (uLongf)expect;
  const int rc = uncompress(out.data() + o, &got, in, (uLong)n);
  if (rc != Z_OK && rc != Z_BUF_ERROR) throw std::runtime_error("TIFF: bad Deflate stream");
  out.resize(o + got);
}

}  // namespace tiff_detail

inline tiff_frame read_tiff(const std::string& name) {
  using namespace tiff_detail;
  std::ifstream f(name, std::ios::binary);
  if (!f) throw std::runtime_error("@read_tiff_image, cannot read tiff file named '" + name + "'");
  const std::vector<uint8_t> b((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  if (b.size() < 8 || !((b[0] == 'I' && b[1] == 'I') || (b[0] == 'M' && b[1] == 'M')))
    throw std::runtime_error("TIFF: not a TIFF file: " + name);
  const Reader r{b, b[0] == 'M'};
  if (r.u16(2) != 42) throw std::runtime_error("TIFF: not a classic TIFF file (BigTIFF is not read): " + name);
  const size_t ifd = r.u32(4);
  const int ntag = r.u16(ifd);
  Entry tags[17];  // the tags below, by slot
  auto slot = [](uint16_t tag) -> int {
    switch (tag) {
      case 256: return 0; case 257: return 1; case 258: return 2; case 259: return 3;
      case 262: return 4; case 273: return 5; case 274: return 6; case 277: return 7;
      case 278: return 8; case 279: return 9; case 284: return 10; case 306: return 11;
      case 317: return 12; case 320: return 13; case 322: return 14; case 338: return 15;
      case 339: return 16;
      default: return -1;
    }
  };
  bool have[17] = {};
  for (int i = 0; i < ntag; i++) {
    const size_t e = ifd + 2 + 12 * (size_t)i;
    const int s = slot(r.u16(e));
    if (s < 0) continue;
    Entry en;
    en.type = r.u16(e + 2);
    en.count = r.u32(e + 4);
    const size_t bytes = type_size(en.type) * en.count;
    en.data = bytes <= 4 ? e + 8 : r.u32(e + 8);
    if (en.data + bytes > b.size()) throw std::runtime_error("TIFF: tag data beyond the file");
    tags[s] = en;
    have[s] = true;
  }
  auto get = [&](int s, uint32_t def) { return have[s] ? value(r, tags[s], 0) : def; };
  if (!have[0] || !have[1] || !have[5]) throw std::runtime_error("TIFF: missing size or strip tags: " + name);
  if (have[14]) throw std::runtime_error("TIFF: tiled files are not read: " + name);
  const uint32_t W = get(0, 0), H = get(1, 0);
  const uint32_t spp = get(7, 1), bps = get(2, 1), comp = get(3, 1), photo = get(4, 1);
  const uint32_t planar = get(10, 1), pred = get(12, 1), orient = get(6, 1), fmt = get(16, 1);
  const uint32_t rps = std::min<uint32_t>(get(8, H), H);
  if (W == 0 || H == 0 || (uint64_t)W * H > (1ull << 31)) throw std::runtime_error("TIFF: bad size");
  if (bps != 8 && bps != 16) throw std::runtime_error("TIFF: only 8- and 16-bit samples are read");
  for (uint32_t s = 1; have[2] && s < tags[2].count; s++)
    if (value(r, tags[2], s) != bps) throw std::runtime_error("TIFF: mixed sample sizes");
  if (fmt != 1) throw std::runtime_error("TIFF: only unsigned integer samples are read");
  if (orient != 1) throw std::runtime_error("TIFF: only top-left orientation is read");
  // libtiff applies a predictor only with the codecs that declare one (LZW, Deflate) and
  // ignores the tag otherwise
  const bool coded = comp == 5 || comp == 8 || comp == 32946;
  if (coded && pred != 1 && pred != 2) throw std::runtime_error("TIFF: unsupported predictor");
  const bool diff = coded && pred == 2;
  if (!(photo <= 3)) throw std::runtime_error("TIFF: unsupported photometric interpretation");
  const uint32_t color = photo == 2 ? 3 : 1;  // samples that carry the color
  if (spp < color) throw std::runtime_error("TIFF: too few samples per pixel");
  if (photo == 2 && spp > 3 && have[15] && value(r, tags[15], 0) == 2)
    throw std::runtime_error("TIFF: unassociated alpha is not read");
  const uint32_t bytes_ps = bps / 8;
  const uint32_t nplanes = planar == 2 ? spp : 1, spp_plane = planar == 2 ? 1 : spp;
  const size_t row_bytes = (size_t)W * spp_plane * bytes_ps;
  const uint32_t strips_per_plane = (H + rps - 1) / rps;
  if (tags[5].count < strips_per_plane * nplanes || !have[9] || tags[9].count < tags[5].count)
    throw std::runtime_error("TIFF: strip tables too short");
  // decode every plane into samples (host byte order)
  std::vector<std::vector<uint16_t>> plane(nplanes, std::vector<uint16_t>((size_t)W * H * spp_plane));
  for (uint32_t p = 0; p < nplanes; p++) {
    for (uint32_t st = 0; st < strips_per_plane; st++) {
      const uint32_t idx = p * strips_per_plane + st;
      const size_t off = value(r, tags[5], idx), cnt = value(r, tags[9], idx);
      if (off + cnt > b.size()) throw std::runtime_error("TIFF: strip beyond the file");
      const uint32_t rows = std::min(rps, H - st * rps);
      const size_t need = row_bytes * rows;
      std::vector<uint8_t> raw;
      raw.reserve(need);
      if (comp == 1) raw.assign(b.begin() + off, b.begin() + off + std::min(cnt, need));
      else if (comp == 5) lzw_decode(b.data() + off, cnt, raw);
      else if (comp == 32773) packbits_decode(b.data() + off, cnt, raw);
      else if (comp == 8 || comp == 32946) deflate_decode(b.data() + off, cnt, need, raw);
      else throw std::runtime_error("TIFF: unsupported compression " + std::to_string(comp));
      if (raw.size() < need) throw std::runtime_error("TIFF: strip shorter than its rows");
      for (uint32_t y = 0; y < rows; y++) {
        uint16_t* dst = plane[p].data() + ((size_t)(st * rps + y) * W) * spp_plane;
        const uint8_t* src = raw.data() + y * row_bytes;
        const size_t ns = (size_t)W * spp_plane;
        for (size_t i = 0; i < ns; i++)
          dst[i] = bytes_ps == 1 ? src[i]
                                 : (r.big ? (uint16_t)(src[2 * i] << 8 | src[2 * i + 1])
                                          : (uint16_t)(src[2 * i + 1] << 8 | src[2 * i]));
        if (diff)  // horizontal differencing, per sample, modulo the sample size
          for (size_t i = spp_plane; i < ns; i++)
            dst[i] = bytes_ps == 1 ? (uint16_t)((dst[i] + dst[i - spp_plane]) & 0xff)
                                   : (uint16_t)(dst[i] + dst[i - spp_plane]);
      }
    }
  }
  auto sample = [&](size_t px, uint32_t s) -> uint32_t {
    return planar == 2 ? plane[s][px] : plane[0][px * spp + s];
  };
  // the 8-bit R, G, B of libtiff's RGBA interface
  std::vector<uint8_t> cmap8;
  if (photo == 3) {
    if (!have[13] || tags[13].count < 3u * (1u << bps) || bps != 8)
      throw std::runtime_error("TIFF: palette images need an 8-bit colormap");
    const uint32_t nc = 1u << bps;
    bool wide = false;
    for (uint32_t i = 0; i < 3 * nc; i++) wide |= value(r, tags[13], i) >= 256;
    cmap8.resize(3 * nc);
    for (uint32_t i = 0; i < 3 * nc; i++) {
      const uint32_t c = value(r, tags[13], i);
      cmap8[i] = (uint8_t)(wide ? c >> 8 : c);
    }
  }
  tiff_frame out;
  out.width = (int)W;
  out.height = (int)H;
  out.gray.resize((size_t)W * H);
  const double fact = 1.0 / 255.0;
  for (size_t px = 0; px < (size_t)W * H; px++) {
    uint32_t R, G, B;
    if (photo == 2) {
      R = sample(px, 0);
      G = sample(px, 1);
      B = sample(px, 2);
      if (bps == 16) {
        R = (R + 128) / 257;
        G = (G + 128) / 257;
        B = (B + 128) / 257;
      }
    } else if (photo == 3) {
      const uint32_t i = sample(px, 0);
      R = cmap8[i];
      G = cmap8[(1u << bps) + i];
      B = cmap8[2 * (1u << bps) + i];
    } else {
      uint32_t v = sample(px, 0);
      if (bps == 16) v >>= 8;
      if (photo == 0) v = 255 - v;
      R = G = B = v;
    }
    double rr = R * fact, gg = G * fact, bb = B * fact;
    rr = rr < 0.0 ? 0.0 : (rr > 1.0 ? 1.0 : rr);
    gg = gg < 0.0 ? 0.0 : (gg > 1.0 ? 1.0 : gg);
    bb = bb < 0.0 ? 0.0 : (bb > 1.0 ? 1.0 : bb);
    out.gray[px] = (uint16_t)(65535.0 * (0.299 * rr + 0.587 * gg + 0.114 * bb));
  }
  if (have[11] && tags[11].type == 2 && tags[11].count > 0) {
    const char* s = reinterpret_cast<const char*>(b.data() + tags[11].data);
    out.dateTime.assign(s, strnlen(s, tags[11].count));
  } else {
    out.dateTime = "dateTime unknown";
  }
  return out;
}

}  // namespace ptgpu
