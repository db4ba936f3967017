// fits.cpp — primary-HDU FITS reader and FLOAT_IMG writer (no CCfits).
//
// Replaces readImage / writeImage of src/fitsUtil.cpp:8-64: BITPIX -32 and -64
// images are accepted (other types are refused with the reference's message),
// BSCALE/BZERO are applied as cfitsio does when reading into doubles, and the
// pixels are narrowed to float exactly as Image::operator[] does
// (include/datatypeUtil.h:158).  Output is a FLOAT_IMG primary HDU that
// overwrites an existing file (the reference's "!" prefix,
// include/datatypeUtil.h:169).  gzip-compressed input is read transparently,
// as cfitsio does.
#include <zlib.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/xbach.h"

namespace {

constexpr size_t kBlock = 2880;
constexpr size_t kCard = 80;

std::string trim(const std::string& s) {
  size_t a = s.find_first_not_of(' ');
  if (a == std::string::npos) return "";
  size_t b = s.find_last_not_of(' ');
  return s.substr(a, b - a + 1);
}

// Value field of a header card: text after "= " up to a comment slash, with
// quotes removed for strings.
std::string card_value(const char* card) {
  std::string v(card + 10, kCard - 10);
  if (!v.empty() && trim(v)[0] == '\'') {
    size_t q0 = v.find('\'');
    size_t q1 = v.find('\'', q0 + 1);
    return q1 == std::string::npos ? v.substr(q0 + 1) : v.substr(q0 + 1, q1 - q0 - 1);
  }
  size_t slash = v.find('/');
  if (slash != std::string::npos) v = v.substr(0, slash);
  return trim(v);
}

uint32_t be32(const unsigned char* p) {
  return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3];
}

uint64_t be64(const unsigned char* p) {
  return ((uint64_t)be32(p) << 32) | be32(p + 4);
}

void put_card(std::string& hdr, const std::string& text) {
  std::string c = text;
  c.resize(kCard, ' ');
  hdr += c;
}

}  // namespace

extern "C" {

int xb_fits_read(const char* path, int* w, int* h, float* out, size_t cap, char* msg,
                 size_t msglen) {
  auto fail = [&](int rc, const std::string& m) {
    if (msg && msglen) std::snprintf(msg, msglen, "%s", m.c_str());
    return rc;
  };
  if (!path || !w || !h) return fail(XB_ERR_ARG, "no path");
  gzFile f = gzopen(path, "rb");
  if (!f) return fail(XB_ERR_IO, std::string("Unable to open file '") + path + "'");
  long bitpix = 0, naxis = -1, n1 = -1, n2 = -1;
  double bscale = 1.0, bzero = 0.0;
  bool end = false;
  std::vector<char> blk(kBlock);
  while (!end) {
    if (gzread(f, blk.data(), (unsigned)kBlock) != (int)kBlock) {
      gzclose(f);
      return fail(XB_ERR_IO, std::string("truncated FITS header in '") + path + "'");
    }
    for (size_t i = 0; i < kBlock / kCard && !end; i++) {
      const char* card = blk.data() + i * kCard;
      std::string key = trim(std::string(card, 8));
      if (key == "END") {
        end = true;
        break;
      }
      if (std::strncmp(card + 8, "= ", 2) != 0) continue;
      std::string v = card_value(card);
      if (key == "BITPIX") bitpix = std::strtol(v.c_str(), nullptr, 10);
      else if (key == "NAXIS") naxis = std::strtol(v.c_str(), nullptr, 10);
      else if (key == "NAXIS1") n1 = std::strtol(v.c_str(), nullptr, 10);
      else if (key == "NAXIS2") n2 = std::strtol(v.c_str(), nullptr, 10);
      else if (key == "BSCALE") bscale = std::strtod(v.c_str(), nullptr);
      else if (key == "BZERO") bzero = std::strtod(v.c_str(), nullptr);
    }
  }
  if (bitpix != -32 && bitpix != -64) {
    gzclose(f);
    return fail(XB_ERR_ARG, "fits image of type" + std::to_string(bitpix) + " is not supported.");
  }
  if (naxis < 2 || n1 <= 0 || n2 <= 0) {
    gzclose(f);
    return fail(XB_ERR_ARG, std::string("'") + path + "' has no 2-D primary image");
  }
  *w = (int)n1;
  *h = (int)n2;
  const size_t n = (size_t)n1 * (size_t)n2;
  if (!out) {  // size query
    gzclose(f);
    return XB_OK;
  }
  if (cap < n) {
    gzclose(f);
    return fail(XB_ERR_CAPACITY, "output buffer too small");
  }
  const size_t bpp = bitpix == -32 ? 4 : 8;
  std::vector<unsigned char> raw(n * bpp);
  size_t got = 0;
  while (got < raw.size()) {
    const unsigned chunk = (unsigned)std::min<size_t>(raw.size() - got, 1u << 30);
    const int r = gzread(f, raw.data() + got, chunk);
    if (r <= 0) break;
    got += (size_t)r;
  }
  gzclose(f);
  if (got != raw.size()) return fail(XB_ERR_IO, std::string("truncated FITS data in '") + path + "'");
  const bool scaled = bscale != 1.0 || bzero != 0.0;
  for (size_t i = 0; i < n; i++) {
    double v;
    if (bpp == 4) {
      const uint32_t u = be32(&raw[i * 4]);
      float fv;
      std::memcpy(&fv, &u, 4);
      v = fv;
    } else {
      const uint64_t u = be64(&raw[i * 8]);
      std::memcpy(&v, &u, 8);
    }
    if (scaled) v = v * bscale + bzero;
    out[i] = (float)v;
  }
  return XB_OK;
}

int xb_fits_write(const char* path, int w, int h, const float* data, char* msg, size_t msglen) {
  return xb_fits_write_cards(path, w, h, data, nullptr, 0, msg, msglen);
}

int xb_fits_write_cards(const char* path, int w, int h, const float* data, const char* cards,
                        int ncards, char* msg, size_t msglen) {
  auto fail = [&](int rc, const std::string& m) {
    if (msg && msglen) std::snprintf(msg, msglen, "%s", m.c_str());
    return rc;
  };
  if (!path || !data || w <= 0 || h <= 0 || ncards < 0 || (ncards > 0 && !cards))
    return fail(XB_ERR_ARG, "bad arguments");
  std::string hdr;
  char buf[96];
  put_card(hdr, "SIMPLE  =                    T / file does conform to FITS standard");
  put_card(hdr, "BITPIX  =                  -32 / number of bits per data pixel");
  put_card(hdr, "NAXIS   =                    2 / number of data axes");
  std::snprintf(buf, sizeof(buf), "NAXIS1  = %20d / length of data axis 1", w);
  put_card(hdr, buf);
  std::snprintf(buf, sizeof(buf), "NAXIS2  = %20d / length of data axis 2", h);
  put_card(hdr, buf);
  put_card(hdr, "EXTEND  =                    T / FITS dataset may contain extensions");
  for (int i = 0; i < ncards; i++) hdr += std::string(cards + (size_t)i * kCard, kCard);
  put_card(hdr, "END");
  hdr.resize((hdr.size() + kBlock - 1) / kBlock * kBlock, ' ');
  FILE* f = std::fopen(path, "wb");  // overwrites, as "!" does for cfitsio
  if (!f) return fail(XB_ERR_IO, std::string("Unable to save file '") + path + "'");
  bool ok = std::fwrite(hdr.data(), 1, hdr.size(), f) == hdr.size();
  const size_t n = (size_t)w * (size_t)h;
  std::vector<unsigned char> body((n * 4 + kBlock - 1) / kBlock * kBlock, 0);
  for (size_t i = 0; i < n; i++) {
    uint32_t u;
    std::memcpy(&u, &data[i], 4);
    body[i * 4 + 0] = (unsigned char)(u >> 24);
    body[i * 4 + 1] = (unsigned char)(u >> 16);
    body[i * 4 + 2] = (unsigned char)(u >> 8);
    body[i * 4 + 3] = (unsigned char)u;
  }
  ok = ok && std::fwrite(body.data(), 1, body.size(), f) == body.size();
  ok = (std::fclose(f) == 0) && ok;
  return ok ? XB_OK : fail(XB_ERR_IO, std::string("Unable to save file '") + path + "'");
}

}  // extern "C"
