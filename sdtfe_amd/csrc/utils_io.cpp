// utils_io.cpp -- host-side helpers of the reference's C surface, exported
// with its prototypes so that its callers link against libsdtfe_amd.so
// alone: particle buffers (src/utils.h:10-12), the forward-hull bitmap
// (src/bitmap.h:9-17) and the map's file output (src/io.h:35-36).  None of
// them touches the GPU.
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/sdtfe_amd.h"

namespace {

// The 8-bit Jet palette of the density image: each channel is 0, 255 or a
// ramp of 64 steps of 255/64 rounded to nearest (half up); red rises over
// indices 96-159 and falls over 224-255, green rises over 32-95 and falls
// over 160-223, blue starts at step 33 of its ramp, holds 32-95 and falls
// over 96-159.
unsigned jet_ramp(int k) { return (unsigned)((255 * k + 32) / 64); }

void jet8(int i, unsigned &r, unsigned &g, unsigned &b) {
  r = i < 96 ? 0 : i < 160 ? jet_ramp(i - 95) : i < 224 ? 255 : jet_ramp(287 - i);
  g = i < 32 ? 0 : i < 96 ? jet_ramp(i - 31) : i < 160 ? 255 : i < 224 ? jet_ramp(223 - i) : 0;
  b = i < 32 ? jet_ramp(i + 33) : i < 96 ? 255 : i < 160 ? jet_ramp(159 - i) : 0;
}

// Little-endian TIFF 6.0 writer for one 8-bit palette image stored as a
// single uncompressed strip.
struct TiffOut {
  std::vector<unsigned char> b;
  void u16(unsigned v) {
    b.push_back((unsigned char)(v & 0xff));
    b.push_back((unsigned char)(v >> 8));
  }
  void u32(uint32_t v) {
    for (int k = 0; k < 4; ++k) b.push_back((unsigned char)(v >> (8 * k)));
  }
  void put_u32(size_t at, uint32_t v) {
    for (int k = 0; k < 4; ++k) b[at + k] = (unsigned char)(v >> (8 * k));
  }
};

}  // namespace

extern "C" {

// src/utils.c:8-16: planar float32 positions (all x, then all y, then all
// z) widened to N x 4 float64 rows (x, y, z, w); w is set to 0 here (the
// reference leaves it unset).  malloc'd: the caller frees it.
double *convert_to_qhdata(const float *rm_buff, size_t n_data) {
  double *q = (double *)std::calloc(n_data > 0 ? 4 * n_data : 4, sizeof(double));
  if (!q) return nullptr;
  for (size_t i = 0; i < n_data; ++i)
    for (size_t j = 0; j < 3; ++j) q[4 * i + j] = (double)rm_buff[j * n_data + i];
  return q;
}

// src/utils.c:18-30: partial Fisher-Yates over the first n_shuffle rows,
// drawing from the C library's rand() as the reference does (the caller
// seeds it with srand): for i = 1 .. n_shuffle - 1, row i - 1 swaps with
// row rand() % (n - i) + i.
void shuffle_buff(double *data, int n_particles_in_buff, int n_shuffle) {
  for (int i = 1; i < n_shuffle; ++i) {
    const int r = std::rand() % (n_particles_in_buff - i) + i;
    double tmp[4];
    std::memcpy(tmp, data + 4 * (size_t)r, sizeof tmp);
    std::memcpy(data + 4 * (size_t)r, data + 4 * (size_t)(i - 1), sizeof tmp);
    std::memcpy(data + 4 * (size_t)(i - 1), tmp, sizeof tmp);
  }
}

// src/utils.c:33-79: rotation of the first n_particles rows about center by
// R = (Qx(theta0) Qy(theta1)) Qz(theta2), each matrix product summed over
// m = 0, 1, 2 from zero, each coordinate as x R[k][0] + y R[k][1] + z R[k][2]
// + center[k] (this file is built without contraction into FMAs, so the
// roundings are the reference's).
void rotate3d(double *data, size_t n_particles, double theta[3], double center[3]) {
  const double cx = std::cos(theta[0]), sx = std::sin(theta[0]);
  const double cy = std::cos(theta[1]), sy = std::sin(theta[1]);
  const double cz = std::cos(theta[2]), sz = std::sin(theta[2]);
  const double qx[3][3] = {{1.0, 0.0, 0.0}, {0.0, cx, -sx}, {0.0, sx, cx}};
  const double qy[3][3] = {{cy, 0.0, sy}, {0.0, 1.0, 0.0}, {-sy, 0.0, cy}};
  const double qz[3][3] = {{cz, -sz, 0.0}, {sz, cz, 0.0}, {0.0, 0.0, 1.0}};
  double p[3][3], r[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double acc = 0.0;
      for (int m = 0; m < 3; ++m) acc += qx[i][m] * qy[m][j];
      p[i][j] = acc;
    }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double acc = 0.0;
      for (int m = 0; m < 3; ++m) acc += p[i][m] * qz[m][j];
      r[i][j] = acc;
    }
  for (size_t i = 0; i < n_particles; ++i) {
    double *d = data + 4 * i;
    const double x = d[0] - center[0], y = d[1] - center[1], z = d[2] - center[2];
    d[0] = x * r[0][0] + y * r[0][1] + z * r[0][2] + center[0];
    d[1] = x * r[1][0] + y * r[1][1] + z * r[1][2] + center[1];
    d[2] = x * r[2][0] + y * r[2][1] + z * r[2][2] + center[2];
  }
}

// src/bitmap.c:8-31.  The reference's words hold 4 bits each (its bits per
// word is sizeof(word) * sizeof(char)), so bit n lives in word n / 4 at
// position n % 4: bit 4 t + f (tet t, face f) is bit f of word t.  The same
// layout is kept so that bitmaps pass between the two libraries.
void set_bit(uint32_t *words, int n) { words[n / 4] |= 1u << (n % 4); }
void clear_bit(uint32_t *words, int n) { words[n / 4] &= ~(1u << (n % 4)); }
int get_bit(uint32_t *words, int n) { return (words[n / 4] >> (n % 4) & 1u) != 0; }
uint32_t *create_bitmap(int n) {
  return (uint32_t *)std::calloc((size_t)(n / 4) + 1, sizeof(uint32_t));
}
void delete_bitmap(uint32_t *words) { std::free(words); }

// src/io.c:89-95: the raw map, rho_size float64 values; returns the count
// written (0 when the file cannot be opened).
int write_rho(char *fname, double **rho, int rho_size) {
  std::FILE *f = std::fopen(fname, "wb");
  if (!f) return 0;
  const size_t n = std::fwrite(*rho, sizeof(double), rho_size > 0 ? (size_t)rho_size : 0, f);
  std::fclose(f);
  return (int)n;
}

// src/io.c:13-86: the map as an 8-bit palette TIFF, row y of the image is
// map row y.  Index of pixel i: 0 where rho <= 0, else
// (unsigned char)(128 (log(rho + 1) - log(min + 1)) / (log(max + 1) -
// log(min + 1))) with max over all pixels (from 0) and min over the
// positive ones, so indices span 0-128 (0 where every positive pixel has one
// value, a 0 / 0 in the reference).  A negative or NaN pixel writes no file.
// The palette is the Jet table above; TIFF palettes hold 16-bit levels, so
// each 8-bit level v is stored as 257 v (the reference hands libtiff its
// 8-bit arrays where 16-bit ones are read).
void write_image(char *file_name, double *rho, int grid_dim) {
  const size_t npix = grid_dim > 0 ? (size_t)grid_dim * (size_t)grid_dim : 0;
  for (size_t i = 0; i < npix; ++i) {
    if (rho[i] < 0.0) {
      std::printf("Negative density!\n");
      std::fflush(stdout);
      return;
    }
    if (rho[i] != rho[i]) {
      std::printf("NaN density!\n");
      std::fflush(stdout);
      return;
    }
  }
  double max_val = 0.0, min_val = DBL_MAX;
  for (size_t i = 0; i < npix; ++i) {
    max_val = rho[i] > max_val ? rho[i] : max_val;
    min_val = (rho[i] > 0.0 && rho[i] < min_val) ? rho[i] : min_val;
  }
  const double max_log = std::log(max_val + 1.0), min_log = std::log(min_val + 1.0);
  const double nf = max_log - min_log;
  std::vector<unsigned char> img(npix, 0);
  for (size_t i = 0; i < npix; ++i)
    if (rho[i] > 0.0) {
      const double v = 128.0 * (std::log(rho[i] + 1.0) - min_log) / nf;
      img[i] = (v >= 0.0 && v < 256.0) ? (unsigned char)v : 0;  // NaN (nf = 0) -> 0
    }

  TiffOut t;
  t.b.insert(t.b.end(), {'I', 'I', 42, 0});
  t.u32(8 + (uint32_t)npix);  // the IFD follows the pixels
  t.b.insert(t.b.end(), img.begin(), img.end());
  if (t.b.size() & 1) t.b.push_back(0);  // IFDs start on a word boundary
  const uint32_t ifd = (uint32_t)t.b.size();
  t.put_u32(4, ifd);
  struct Tag {
    unsigned id, type;
    uint32_t count, value;
  };
  constexpr unsigned kTags = 12;
  const uint32_t cmap_at = ifd + 2 + 12 * kTags + 4;
  const Tag tags[] = {
      {256, 4, 1, (uint32_t)grid_dim},  // ImageWidth
      {257, 4, 1, (uint32_t)grid_dim},  // ImageLength
      {258, 3, 1, 8},                   // BitsPerSample
      {259, 3, 1, 1},                   // Compression: none
      {262, 3, 1, 3},                   // PhotometricInterpretation: palette
      {273, 4, 1, 8},                   // StripOffsets
      {274, 3, 1, 1},                   // Orientation: top left
      {277, 3, 1, 1},                   // SamplesPerPixel
      {278, 4, 1, (uint32_t)grid_dim},  // RowsPerStrip: one strip
      {279, 4, 1, (uint32_t)npix},      // StripByteCounts
      {284, 3, 1, 1},                   // PlanarConfiguration: contiguous
      {320, 3, 3 * 256, cmap_at},       // ColorMap
  };
  static_assert(sizeof(tags) / sizeof(tags[0]) == kTags, "tag count");
  t.u16(kTags);
  for (const Tag &g : tags) {
    t.u16(g.id);
    t.u16(g.type);
    t.u32(g.count);
    if (g.type == 3 && g.count == 1) {
      t.u16(g.value);
      t.u16(0);
    } else {
      t.u32(g.value);
    }
  }
  t.u32(0);  // no further IFD
  unsigned rgb[3][256];
  for (int i = 0; i < 256; ++i) jet8(i, rgb[0][i], rgb[1][i], rgb[2][i]);
  for (int c = 0; c < 3; ++c)
    for (int i = 0; i < 256; ++i) t.u16(257u * rgb[c][i]);

  std::FILE *f = std::fopen(file_name, "wb");
  if (!f) {
    std::printf("Cannot open %s\n", file_name);
    return;
  }
  std::fwrite(t.b.data(), 1, t.b.size(), f);
  std::fclose(f);
}

}  // extern "C"
