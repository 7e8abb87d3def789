// Per-pixel feature functors: the HIP/host replacement of the reference's
// user-supplied OpenCL hook `feat_t computeFeature(image, nChannels, width,
// height, coords, treeLeftChildren, treePosteriors, imageNodesID, features,
// featDim)` (test/feature.cl:29-34; contract in src/padenti/padenti.hpp.in:
// 170-175).  A functor is a struct with
//   using ImgType, FeatType; static constexpr nChannels, FeatDim;
//   static FeatType eval(const ImageView<ImgType,nChannels>&, int x, int y,
//                        const FeatType* params);
// compiled for host and device.  Reads outside the image follow the
// reference sampler (CLK_ADDRESS_CLAMP + an explicit border test in the hook):
// each functor returns its own background value there.
//
// A response goes to the left child when response <= threshold
// (hist_update.cl:154, predict.cl:75).
#pragma once
#include <math.h>
#include "pd_math.hpp"

#pragma clang fp contract(off)

namespace pd {

// Interleaved row-major image (the reference Image<type,nChannels> layout
// for nChannels <= 4: image.hpp:34-111, uploaded as image2d_t).
template <typename ImgT, int NCH>
struct ImageView {
  const ImgT *data;
  int w, h;
  PD_HD ImgT at(int x, int y, int c) const
  {
    return data[((uint32_t)y * (uint32_t)w + (uint32_t)x) * (uint32_t)NCH + (uint32_t)c];
  }
  PD_HD bool inside(int x, int y) const { return x >= 0 && x < w && y >= 0 && y < h; }
};

// round(q) half away from zero (OpenCL round(), test/feature.cl:50); a
// quotient whose magnitude is not below 2^20 counts as leaving the image,
// which keeps the float->int conversion defined on every target.
PD_HD bool offset_ok(float q) { return fabsf(q) < 1048576.0f; }

// C1: float depth in metres.  f = I(x + round(u/d(x))) - I(x + round(v/d(x))),
// p = (ux, uy, vx, vy) in pixel*m, out-of-image reads 10 m, d(x) <= 0 -> 0.
struct DepthF32Feature {
  using ImgType = float;
  using FeatType = float;
  static constexpr int nChannels = 1;
  static constexpr int FeatDim = 4;
  static constexpr int kind = 0;
  PD_HD static float probe(const ImageView<float, 1> &im, int x, int y, float d, float px, float py)
  {
    float qx = roundf(px / d), qy = roundf(py / d);
    if (!offset_ok(qx) || !offset_ok(qy)) return 10.0f;
    int ox = x + (int)qx, oy = y + (int)qy;
    return im.inside(ox, oy) ? im.at(ox, oy, 0) : 10.0f;
  }
  PD_HD static float eval(const ImageView<float, 1> &im, int x, int y, const float *p)
  {
    float d = im.at(x, y, 0);
    if (!(d > 0.0f)) return 0.0f;
    float a = probe(im, x, y, d, p[0], p[1]);
    float b = probe(im, x, y, d, p[2], p[3]);
    return a - b;
  }
};

// C2/C3: uint16 depth in millimetres, float response; p in pixel*mm,
// out-of-image reads 10000 mm, d(x) == 0 -> 0.
struct DepthU16Feature {
  using ImgType = uint16_t;
  using FeatType = float;
  static constexpr int nChannels = 1;
  static constexpr int FeatDim = 4;
  static constexpr int kind = 1;
  PD_HD static float probe(const ImageView<uint16_t, 1> &im, int x, int y, float d, float px, float py)
  {
    float qx = roundf(px / d), qy = roundf(py / d);
    if (!offset_ok(qx) || !offset_ok(qy)) return 10000.0f;
    int ox = x + (int)qx, oy = y + (int)qy;
    return im.inside(ox, oy) ? (float)im.at(ox, oy, 0) : 10000.0f;
  }
  PD_HD static float eval(const ImageView<uint16_t, 1> &im, int x, int y, const float *p)
  {
    float d = (float)im.at(x, y, 0);
    if (!(d > 0.0f)) return 0.0f;
    float a = probe(im, x, y, d, p[0], p[1]);
    float b = probe(im, x, y, d, p[2], p[3]);
    return a - b;
  }
};

// C4: RGB-D frames as uint16 x 4 (R, G, B, depth mm).  Depth-normalised
// offset channel difference over a channel pair:
// f = I_ca(x + round(u/d)) - I_cb(x + round(v/d)), p = (ux, uy, vx, vy, ca, cb),
// channel = (int)c clamped to [0, 3], out-of-image reads 10000.
struct RGBDFeature {
  using ImgType = uint16_t;
  using FeatType = float;
  static constexpr int nChannels = 4;
  static constexpr int FeatDim = 6;
  static constexpr int kind = 2;
  PD_HD static int chan(float c)
  {
    if (!(c > 0.0f)) return 0;
    if (c >= 3.0f) return 3;
    return (int)c;
  }
  PD_HD static float probe(const ImageView<uint16_t, 4> &im, int x, int y, float d, float px, float py,
                           int ch)
  {
    float qx = roundf(px / d), qy = roundf(py / d);
    if (!offset_ok(qx) || !offset_ok(qy)) return 10000.0f;
    int ox = x + (int)qx, oy = y + (int)qy;
    return im.inside(ox, oy) ? (float)im.at(ox, oy, ch) : 10000.0f;
  }
  PD_HD static float eval(const ImageView<uint16_t, 4> &im, int x, int y, const float *p)
  {
    float d = (float)im.at(x, y, 3);
    if (!(d > 0.0f)) return 0.0f;
    float a = probe(im, x, y, d, p[0], p[1], chan(p[4]));
    float b = probe(im, x, y, d, p[2], p[3], chan(p[5]));
    return a - b;
  }
};

// The reference's own demo hook (test/feature.cl:29-64) with feat_t = short,
// FeatDim 2: depth = (short)I(x); offset = x + (int)round((float)u*500/depth);
// out-of-image -> 10000; response = depth - (d2 ? d2 : 10000) in short.
// depth == 0 (a division by zero in the reference) gives 0.
struct DemoS16Feature {
  using ImgType = uint16_t;
  using FeatType = int16_t;
  static constexpr int nChannels = 1;
  static constexpr int FeatDim = 2;
  static constexpr int kind = 3;
  PD_HD static int16_t eval(const ImageView<uint16_t, 1> &im, int x, int y, const int16_t *p)
  {
    int16_t depth = (int16_t)im.at(x, y, 0);
    if (depth == 0) return 0;
    float fd = (float)depth;
    float qx = roundf((float)p[0] * 500.0f / fd);
    float qy = roundf((float)p[1] * 500.0f / fd);
    int16_t d2;
    if (!offset_ok(qx) || !offset_ok(qy)) {
      d2 = (int16_t)10000;
    } else {
      int ox = x + (int)qx, oy = y + (int)qy;
      d2 = im.inside(ox, oy) ? (int16_t)im.at(ox, oy, 0) : (int16_t)10000;
    }
    return (int16_t)(depth - (d2 ? d2 : (int16_t)10000));
  }
};

// Threshold padding value: compares "left" against every response, so a
// padded slot never counts as "right" for a finite response.
template <typename T>
PD_HD T pad_threshold();
template <>
PD_HD float pad_threshold<float>() { return __builtin_inff(); }
template <>
PD_HD int16_t pad_threshold<int16_t>() { return (int16_t)32767; }

}  // namespace pd
