// image_io.cpp — the reference's image output for a rendered W*H*3 float
// image in its storage order (row 0 = top, render.h:1234):
//   to_int  int(pow(clamp(x, 0, 1), 1/2.2) * 255 + 0.5)   (ppm.h:14-27)
//   PPM     "P3\n<W> <H>\n255\n" then "r g b " per pixel in storage order
//           (save_ppm_file, ppm.h:29-35)
//   BMP     54-byte header, 24-bit BGR rows bottom-up padded to 4 bytes, the
//           first file row being image row H-1 (Render::OutputBmp
//           render.h:1507-1560 over BitMap::Write_Bmp bitmap.cpp:190-285);
//           gamma_offset > 0.01 in magnitude replaces 1/2.2 (render.h:1540).
//   HDR     Radiance RGBE (Render::OutputHdr render.h:1647-1667 over
//           HDRImage::save hdr.h:212-246): a five-line header, "-Y H +X W",
//           then per scanline (image rows 0..H-1) the marker 2 2 W_hi W_lo
//           and the R, G, B, E planes as literal runs of at most 127 bytes;
//           a pixel is (r, g, b) * m * 256 / d truncated to bytes with
//           d = max(r, g, b) = m * 2^e (frexp) and E = e + 128, or all zero
//           when d <= 1e-32.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/prt.h"
#include "prt_internal.h"

using namespace prt;

namespace {

const double kInvGamma = 0.4545454545454545454545454545454545454545;  // C_1_22

inline double clamp01(double x) {
  if (x < 0.0) return 0.0;
  if (x > 1.0) return 1.0;
  return x;
}
inline int to_int(double x, double g) { return int(std::pow(clamp01(x), g) * 255 + 0.5); }

int check_image(const float* rgb, int w, int h) {
  if (!rgb || w <= 0 || h <= 0) return fail(PRT_E_INVALID, "bad image arguments");
  return PRT_OK;
}

void put32(unsigned char* p, uint32_t v) {
  p[0] = (unsigned char)v;
  p[1] = (unsigned char)(v >> 8);
  p[2] = (unsigned char)(v >> 16);
  p[3] = (unsigned char)(v >> 24);
}

}  // namespace

extern "C" {

int prt_image_to_rgb8(const float* rgb, int width, int height, double gamma, uint8_t* out) {
  int rc;
  if ((rc = check_image(rgb, width, height))) return rc;
  if (!out) return fail(PRT_E_INVALID, "null output");
  const double g = gamma > 0.0 ? gamma : kInvGamma;
  const size_t n = (size_t)width * height * 3;
  for (size_t i = 0; i < n; ++i) out[i] = (uint8_t)to_int(rgb[i], g);
  return PRT_OK;
}

int prt_write_ppm(const char* path, const float* rgb, int width, int height) {
  int rc;
  if ((rc = check_image(rgb, width, height))) return rc;
  FILE* f = path ? std::fopen(path, "wb") : nullptr;
  if (!f) return fail(PRT_E_INVALID, std::string("cannot write ") + (path ? path : "(null)"));
  std::fprintf(f, "P3\n%d %d\n%d\n", width, height, 255);
  const size_t n = (size_t)width * height;
  for (size_t i = 0; i < n; ++i)
    std::fprintf(f, "%d %d %d ", to_int(rgb[3 * i], kInvGamma), to_int(rgb[3 * i + 1], kInvGamma),
                 to_int(rgb[3 * i + 2], kInvGamma));
  const bool ok = std::ferror(f) == 0;
  std::fclose(f);
  return ok ? PRT_OK : fail(PRT_E_INVALID, std::string("write error on ") + path);
}

int prt_write_bmp(const char* path, const float* rgb, int width, int height, double gamma_offset) {
  int rc;
  if ((rc = check_image(rgb, width, height))) return rc;
  const double g = std::fabs(gamma_offset) > 0.01 ? gamma_offset : kInvGamma;
  const int row = width * 3 + width % 4;  // = 4-byte padded row
  const uint32_t data = (uint32_t)row * (uint32_t)height;
  unsigned char hdr[54];
  std::memset(hdr, 0, sizeof(hdr));
  hdr[0] = 'B';
  hdr[1] = 'M';
  put32(hdr + 2, 54 + data);
  put32(hdr + 10, 54);
  put32(hdr + 14, 40);
  put32(hdr + 18, (uint32_t)width);
  put32(hdr + 22, (uint32_t)height);
  hdr[26] = 1;   // planes
  hdr[28] = 24;  // bits per pixel
  put32(hdr + 34, data);
  put32(hdr + 38, 1);  // pixels per metre, as the reference writes them
  put32(hdr + 42, 1);
  FILE* f = path ? std::fopen(path, "wb") : nullptr;
  if (!f) return fail(PRT_E_INVALID, std::string("cannot write ") + (path ? path : "(null)"));
  std::fwrite(hdr, 1, sizeof(hdr), f);
  std::vector<unsigned char> line((size_t)row, 0);
  for (int i = 0; i < height; ++i) {
    const float* src = rgb + (size_t)(height - i - 1) * width * 3;
    for (int j = 0; j < width; ++j) {
      line[3 * j + 0] = (unsigned char)to_int(src[3 * j + 2], g);
      line[3 * j + 1] = (unsigned char)to_int(src[3 * j + 1], g);
      line[3 * j + 2] = (unsigned char)to_int(src[3 * j + 0], g);
    }
    std::fwrite(line.data(), 1, line.size(), f);
  }
  const bool ok = std::ferror(f) == 0;
  std::fclose(f);
  return ok ? PRT_OK : fail(PRT_E_INVALID, std::string("write error on ") + path);
}

int prt_write_hdr(const char* path, const float* rgb, int width, int height) {
  int rc;
  if ((rc = check_image(rgb, width, height))) return rc;
  if (width > 0xFFFF) return fail(PRT_E_INVALID, "HDR scanlines hold at most 65535 pixels");
  FILE* f = path ? std::fopen(path, "wb") : nullptr;
  if (!f) return fail(PRT_E_INVALID, std::string("cannot write ") + (path ? path : "(null)"));
  std::fputs("#?RADIANCE\n# Made with 100% pure HDR Shop\nFORMAT=32-bit_rle_rgbe\nEXPOSURE=1.0000000000000\n\n", f);
  std::fprintf(f, "-Y %d +X %d\n", height, width);
  std::vector<unsigned char> planes((size_t)4 * width);
  std::vector<unsigned char> out;
  for (int y = 0; y < height; ++y) {
    const float* src = rgb + (size_t)y * width * 3;
    for (int x = 0; x < width; ++x) {
      const double r = src[3 * x], g = src[3 * x + 1], b = src[3 * x + 2];
      const double d = std::max(r, std::max(g, b));
      unsigned char q[4] = {0, 0, 0, 0};
      if (d > 1e-32) {
        int e;
        const double m = std::frexp(d, &e);
        const double k = m * 256.0 / d;
        q[0] = (unsigned char)(r * k);
        q[1] = (unsigned char)(g * k);
        q[2] = (unsigned char)(b * k);
        q[3] = (unsigned char)(e + 128);
      }
      for (int c = 0; c < 4; ++c) planes[(size_t)c * width + x] = q[c];
    }
    out.assign({0x02, 0x02, (unsigned char)((width >> 8) & 0xFF), (unsigned char)(width & 0xFF)});
    for (int c = 0; c < 4; ++c)
      for (int x = 0; x < width;) {
        const int run = std::min(127, width - x);
        out.push_back((unsigned char)run);
        out.insert(out.end(), planes.begin() + (size_t)c * width + x, planes.begin() + (size_t)c * width + x + run);
        x += run;
      }
    std::fwrite(out.data(), 1, out.size(), f);
  }
  const bool ok = std::ferror(f) == 0;
  std::fclose(f);
  return ok ? PRT_OK : fail(PRT_E_INVALID, std::string("write error on ") + path);
}

}  // extern "C"
