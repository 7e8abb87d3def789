// Host-side image I/O of the ingestion path: a PNG decoder / encoder standing
// in for cv::imread(path, -1) and cv::imwrite (the reference's loaders,
// src/padenti/cv_image_loader_impl.hpp:36-133), and the directory pairing of
// the TrainingSet constructor (src/padenti/training_set_impl.hpp:52-74).
// Implements ISO/IEC 15948 (PNG): chunk CRCs, zlib streams (zlib), the five
// scanline filters, Adam7 interlacing, bit depths 1-16 and all five colour
// types.  Channel order follows OpenCV: B, G, R(, A).
#include "padenti/pd_runtime.hpp"

#include <dirent.h>
#include <sys/stat.h>
#include <zlib.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

using pd::engine::guarded;

namespace {

void need(bool ok, const std::string &msg)
{
  if (!ok) throw pd::engine::ArgError(msg);
}

uint32_t be32(const uint8_t *p) { return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3]; }
void put32(std::vector<uint8_t> &v, uint32_t x)
{
  v.push_back((uint8_t)(x >> 24));
  v.push_back((uint8_t)(x >> 16));
  v.push_back((uint8_t)(x >> 8));
  v.push_back((uint8_t)x);
}

const uint8_t SIG[8] = {137, 80, 78, 71, 13, 10, 26, 10};

struct Png {
  uint32_t w = 0, h = 0, depth = 0, ctype = 0, interlace = 0;
  std::vector<uint8_t> plte, trns, idat;
  uint32_t in_ch() const { return ctype == 0 ? 1 : ctype == 2 ? 3 : ctype == 3 ? 1 : ctype == 4 ? 2 : 4; }
  uint32_t out_ch() const
  {
    if (ctype == 0) return 1;
    if (ctype == 2 || ctype == 3) return trns.empty() ? 3 : 4;
    return 4;
  }
  uint32_t out_depth() const { return depth == 16 ? 16 : 8; }
};

std::vector<uint8_t> read_file(const char *path)
{
  FILE *f = std::fopen(path, "rb");
  need(f != nullptr, std::string("cannot open ") + path);
  std::vector<uint8_t> v;
  uint8_t buf[1 << 16];
  size_t k;
  while ((k = std::fread(buf, 1, sizeof(buf), f)) > 0) v.insert(v.end(), buf, buf + k);
  std::fclose(f);
  return v;
}

Png parse(const std::vector<uint8_t> &f, bool header_only)
{
  need(f.size() >= 8 && std::memcmp(f.data(), SIG, 8) == 0, "not a PNG file");
  Png p;
  size_t o = 8;
  bool ihdr = false, iend = false;
  while (o + 12 <= f.size()) {
    const uint32_t len = be32(&f[o]);
    need(len <= f.size() - o - 12, "truncated PNG chunk");
    const uint8_t *type = &f[o + 4], *data = &f[o + 8];
    const uint32_t crc = be32(&f[o + 8 + len]);
    need((uint32_t)crc32(0u, type, len + 4) == crc, "PNG chunk CRC mismatch");
    const std::string t((const char *)type, 4);
    if (t == "IHDR") {
      need(len == 13 && !ihdr, "bad IHDR");
      p.w = be32(data);
      p.h = be32(data + 4);
      p.depth = data[8];
      p.ctype = data[9];
      need(data[10] == 0 && data[11] == 0, "unknown PNG compression or filter method");
      p.interlace = data[12];
      need(p.interlace <= 1, "unknown PNG interlace method");
      need(p.w >= 1 && p.h >= 1 && p.w < (1u << 31) && p.h < (1u << 31), "bad PNG size");
      const uint32_t d = p.depth, c = p.ctype;
      const bool ok = (c == 0 && (d == 1 || d == 2 || d == 4 || d == 8 || d == 16)) ||
                      (c == 3 && (d == 1 || d == 2 || d == 4 || d == 8)) ||
                      ((c == 2 || c == 4 || c == 6) && (d == 8 || d == 16));
      need(ok, "bad PNG bit depth / colour type");
      ihdr = true;
    } else {
      need(ihdr, "PNG chunk before IHDR");
      if (t == "PLTE") {
        need(len % 3 == 0 && len <= 768, "bad PLTE");
        p.plte.assign(data, data + len);
      } else if (t == "tRNS") {
        p.trns.assign(data, data + len);
      } else if (t == "IDAT") {
        if (!header_only) p.idat.insert(p.idat.end(), data, data + len);  // tRNS still decides the channels
      } else if (t == "IEND") {
        iend = true;
        break;
      } else {
        need((type[0] & 0x20) != 0, "unknown critical PNG chunk " + t);
      }
    }
    o += 12 + (size_t)len;
  }
  need(ihdr, "PNG without IHDR");
  need(iend, "PNG without IEND");
  if (p.ctype == 3) need(!p.plte.empty(), "palette PNG without PLTE");
  return p;
}

uint8_t paeth(int a, int b, int c)
{
  const int pp = a + b - c, pa = std::abs(pp - a), pb = std::abs(pp - b), pc = std::abs(pp - c);
  if (pa <= pb && pa <= pc) return (uint8_t)a;
  if (pb <= pc) return (uint8_t)b;
  return (uint8_t)c;
}

// Undo the scanline filters of one (sub)image in place; returns the byte
// position after it.  Rows are 1 filter byte + rb bytes.
size_t unfilter(uint8_t *d, size_t avail, uint32_t rows, size_t rb, uint32_t bpp)
{
  need(avail >= (size_t)rows * (rb + 1), "PNG image data too short");
  const uint8_t *prev = nullptr;
  for (uint32_t y = 0; y < rows; y++) {
    uint8_t *row = d + (size_t)y * (rb + 1);
    const uint8_t ft = row[0];
    uint8_t *x = row + 1;
    switch (ft) {
      case 0: break;
      case 1:
        for (size_t i = bpp; i < rb; i++) x[i] = (uint8_t)(x[i] + x[i - bpp]);
        break;
      case 2:
        if (prev)
          for (size_t i = 0; i < rb; i++) x[i] = (uint8_t)(x[i] + prev[i]);
        break;
      case 3:
        for (size_t i = 0; i < rb; i++) {
          const int a = i >= bpp ? x[i - bpp] : 0, b = prev ? prev[i] : 0;
          x[i] = (uint8_t)(x[i] + ((a + b) >> 1));
        }
        break;
      case 4:
        for (size_t i = 0; i < rb; i++) {
          const int a = i >= bpp ? x[i - bpp] : 0, b = prev ? prev[i] : 0, c = (prev && i >= bpp) ? prev[i - bpp] : 0;
          x[i] = (uint8_t)(x[i] + paeth(a, b, c));
        }
        break;
      default: need(false, "bad PNG filter type");
    }
    prev = x;
  }
  return (size_t)rows * (rb + 1);
}

// sample k (0-based) of a filtered row at bit depth d
inline uint32_t sample_at(const uint8_t *x, size_t k, uint32_t d)
{
  if (d == 8) return x[k];
  if (d == 16) return ((uint32_t)x[2 * k] << 8) | x[2 * k + 1];
  const size_t bit = k * d;
  return (x[bit >> 3] >> (8 - d - (bit & 7))) & ((1u << d) - 1u);
}

// decoded pixels: out_ch channels of out_depth bits (B, G, R, A order)
void decode(const Png &p, uint8_t *out)
{
  const uint32_t ich = p.in_ch(), och = p.out_ch(), d = p.depth;
  const uint32_t bpp = std::max<uint32_t>(1, ich * d / 8);
  struct Pass { uint32_t x0, y0, dx, dy; };
  const Pass adam7[7] = {{0, 0, 8, 8}, {4, 0, 8, 8}, {0, 4, 4, 8}, {2, 0, 4, 4}, {0, 2, 2, 4}, {1, 0, 2, 2}, {0, 1, 1, 2}};
  const Pass whole[1] = {{0, 0, 1, 1}};
  const Pass *passes = p.interlace ? adam7 : whole;
  const int np = p.interlace ? 7 : 1;
  // inflated size
  size_t total = 0;
  for (int k = 0; k < np; k++) {
    const Pass &q = passes[k];
    if (p.w <= q.x0 || p.h <= q.y0) continue;
    const uint64_t pw = (p.w - q.x0 + q.dx - 1) / q.dx, ph = (p.h - q.y0 + q.dy - 1) / q.dy;
    const uint64_t rb = (pw * ich * d + 7) / 8;
    total += (size_t)(ph * (rb + 1));
  }
  std::vector<uint8_t> raw(total + 1);
  z_stream zs;
  std::memset(&zs, 0, sizeof(zs));
  need(inflateInit(&zs) == Z_OK, "zlib inflateInit failed");
  zs.next_in = const_cast<Bytef *>(p.idat.data());
  zs.avail_in = (uInt)p.idat.size();
  zs.next_out = raw.data();
  zs.avail_out = (uInt)raw.size();
  const int rc = inflate(&zs, Z_FINISH);
  const size_t got = raw.size() - zs.avail_out;
  inflateEnd(&zs);
  need(rc == Z_STREAM_END, "corrupt PNG image data (zlib)");
  need(got == total, "PNG image data has the wrong size");
  const uint32_t maxv = (1u << d) - 1u;
  const uint32_t ob = p.out_depth() / 8;
  // tRNS keys
  uint32_t key_g = 0xFFFFFFFFu, key_r = 0, key_gg = 0, key_b = 0;
  bool key_rgb = false;
  if (p.ctype == 0 && p.trns.size() >= 2) key_g = ((uint32_t)p.trns[0] << 8) | p.trns[1];
  if (p.ctype == 2 && p.trns.size() >= 6) {
    key_rgb = true;
    key_r = ((uint32_t)p.trns[0] << 8) | p.trns[1];
    key_gg = ((uint32_t)p.trns[2] << 8) | p.trns[3];
    key_b = ((uint32_t)p.trns[4] << 8) | p.trns[5];
  }
  (void)key_g;
  size_t pos = 0;
  for (int k = 0; k < np; k++) {
    const Pass &q = passes[k];
    if (p.w <= q.x0 || p.h <= q.y0) continue;
    const uint32_t pw = (p.w - q.x0 + q.dx - 1) / q.dx, ph = (p.h - q.y0 + q.dy - 1) / q.dy;
    const size_t rb = ((size_t)pw * ich * d + 7) / 8;
    uint8_t *base = raw.data() + pos;
    pos += unfilter(base, total - pos, ph, rb, bpp);
    for (uint32_t yy = 0; yy < ph; yy++) {
      const uint8_t *x = base + (size_t)yy * (rb + 1) + 1;
      const uint32_t y = q.y0 + yy * q.dy;
      for (uint32_t xx = 0; xx < pw; xx++) {
        const uint32_t X = q.x0 + xx * q.dx;
        uint32_t v[4] = {0, 0, 0, 0};  // B, G, R, A in output scale
        if (p.ctype == 0) {
          uint32_t g = sample_at(x, xx, d);
          if (d < 8) g = g * (255u / maxv);
          v[0] = g;
        } else if (p.ctype == 3) {
          const uint32_t idx = sample_at(x, xx, d);
          need(3 * idx + 2 < p.plte.size(), "PNG palette index out of range");
          v[0] = p.plte[3 * idx + 2];
          v[1] = p.plte[3 * idx + 1];
          v[2] = p.plte[3 * idx];
          v[3] = idx < p.trns.size() ? p.trns[idx] : 255u;
        } else if (p.ctype == 2) {
          const uint32_t r = sample_at(x, 3 * (size_t)xx, d), g = sample_at(x, 3 * (size_t)xx + 1, d),
                         b = sample_at(x, 3 * (size_t)xx + 2, d);
          v[0] = b;
          v[1] = g;
          v[2] = r;
          v[3] = (key_rgb && r == key_r && g == key_gg && b == key_b) ? 0u : maxv;
        } else if (p.ctype == 4) {
          const uint32_t g = sample_at(x, 2 * (size_t)xx, d);
          v[0] = v[1] = v[2] = g;
          v[3] = sample_at(x, 2 * (size_t)xx + 1, d);
        } else {
          v[0] = sample_at(x, 4 * (size_t)xx + 2, d);
          v[1] = sample_at(x, 4 * (size_t)xx + 1, d);
          v[2] = sample_at(x, 4 * (size_t)xx, d);
          v[3] = sample_at(x, 4 * (size_t)xx + 3, d);
        }
        uint8_t *o = out + ((size_t)y * p.w + X) * och * ob;
        for (uint32_t c = 0; c < och; c++) {
          if (ob == 1) {
            o[c] = (uint8_t)v[c];
          } else {
            const uint16_t s = (uint16_t)v[c];
            std::memcpy(o + 2 * c, &s, 2);
          }
        }
      }
    }
  }
}

void chunk(std::vector<uint8_t> &f, const char *type, const uint8_t *data, size_t len)
{
  put32(f, (uint32_t)len);
  const size_t o = f.size();
  f.insert(f.end(), type, type + 4);
  if (len) f.insert(f.end(), data, data + len);
  put32(f, (uint32_t)crc32(0u, f.data() + o, (uInt)(len + 4)));
}

void write_png(const char *path, uint32_t w, uint32_t h, uint32_t ch, uint32_t depth, const uint8_t *px)
{
  need(w >= 1 && h >= 1 && w < (1u << 31) && h < (1u << 31), "bad PNG size");
  need(ch == 1 || ch == 3 || ch == 4, "channels must be 1, 3 or 4");
  need(depth == 8 || depth == 16, "bit depth must be 8 or 16");
  const uint32_t ob = depth / 8;
  const size_t rb = (size_t)w * ch * ob;
  std::vector<uint8_t> raw((rb + 1) * h);
  for (uint32_t y = 0; y < h; y++) {
    uint8_t *r = raw.data() + (size_t)y * (rb + 1);
    r[0] = 0;
    for (uint32_t x = 0; x < w; x++)
      for (uint32_t c = 0; c < ch; c++) {
        // file order R, G, B(, A) from B, G, R(, A)
        const uint32_t src = (ch >= 3 && c < 3) ? 2 - c : c;
        const uint8_t *s = px + (((size_t)y * w + x) * ch + src) * ob;
        uint8_t *dd = r + 1 + ((size_t)x * ch + c) * ob;
        if (ob == 1) {
          dd[0] = s[0];
        } else {
          uint16_t v;
          std::memcpy(&v, s, 2);
          dd[0] = (uint8_t)(v >> 8);
          dd[1] = (uint8_t)v;
        }
      }
  }
  uLongf zlen = compressBound((uLong)raw.size());
  std::vector<uint8_t> z(zlen);
  need(compress2(z.data(), &zlen, raw.data(), (uLong)raw.size(), 6) == Z_OK, "zlib compress failed");
  std::vector<uint8_t> f(SIG, SIG + 8);
  uint8_t ih[13];
  const uint32_t ct = ch == 1 ? 0 : ch == 3 ? 2 : 6;
  const uint32_t v4[2] = {w, h};
  for (int i = 0; i < 2; i++)
    for (int b = 0; b < 4; b++) ih[4 * i + b] = (uint8_t)(v4[i] >> (24 - 8 * b));
  ih[8] = (uint8_t)depth;
  ih[9] = (uint8_t)ct;
  ih[10] = ih[11] = ih[12] = 0;
  chunk(f, "IHDR", ih, 13);
  chunk(f, "IDAT", z.data(), zlen);
  chunk(f, "IEND", nullptr, 0);
  FILE *out = std::fopen(path, "wb");
  need(out != nullptr, std::string("cannot write ") + path);
  const size_t k = std::fwrite(f.data(), 1, f.size(), out);
  const int cl = std::fclose(out);
  need(k == f.size() && cl == 0, std::string("write failed: ") + path);
}

bool ends_with(const std::string &s, const std::string &suf)
{
  return s.size() >= suf.size() && s.compare(s.size() - suf.size(), suf.size(), suf) == 0;
}

}  // namespace

extern "C" {

int pd_png_read(const char *path, pd_png_info *info, void *pixels, uint64_t capacity)
{
  return guarded([&] {
    need(path && info, "null argument");
    const std::vector<uint8_t> f = read_file(path);
    const Png p = parse(f, pixels == nullptr);
    info->width = p.w;
    info->height = p.h;
    info->channels = p.out_ch();
    info->bit_depth = p.out_depth();
    if (!pixels) return;
    const uint64_t bytes = (uint64_t)p.w * p.h * p.out_ch() * (p.out_depth() / 8);
    need(capacity >= bytes, "pixel buffer too small");
    decode(p, (uint8_t *)pixels);
  });
}

int pd_png_write(const char *path, uint32_t width, uint32_t height, uint32_t channels, uint32_t bit_depth,
                 const void *pixels)
{
  return guarded([&] {
    need(path && pixels, "null argument");
    write_png(path, width, height, channels, bit_depth, (const uint8_t *)pixels);
  });
}

int pd_list_pairs(const char *dir, const char *data_suffix, const char *labels_suffix, char *buf, uint64_t cap,
                  uint32_t *n_pairs, uint64_t *needed)
{
  return guarded([&] {
    need(dir && data_suffix && labels_suffix && n_pairs && needed, "null argument");
    DIR *d = opendir(dir);
    need(d != nullptr, std::string("cannot open directory ") + dir);
    std::vector<std::string> names;
    while (dirent *e = readdir(d)) names.push_back(e->d_name);
    closedir(d);
    std::sort(names.begin(), names.end());
    const std::string root(dir), ds(data_suffix), ls(labels_suffix);
    std::vector<std::string> out;
    for (const std::string &nm : names) {
      const std::string full = root + "/" + nm;
      struct stat st;
      if (stat(full.c_str(), &st) != 0 || S_ISDIR(st.st_mode)) continue;
      if (!ends_with(nm, ds)) continue;
      const std::string lab = root + "/" + nm.substr(0, nm.size() - ds.size()) + ls;
      struct stat sl;
      if (stat(lab.c_str(), &sl) != 0) continue;
      out.push_back(full);
      out.push_back(lab);
    }
    uint64_t n = 0;
    for (const std::string &s : out) n += s.size() + 1;
    *needed = n;
    *n_pairs = (uint32_t)(out.size() / 2);
    if (cap == 0 || buf == nullptr) return;
    need(cap >= n, "buffer too small for the file list");
    char *o = buf;
    for (const std::string &s : out) {
      std::memcpy(o, s.c_str(), s.size() + 1);
      o += s.size() + 1;
    }
  });
}

}  // extern "C"
