// Per-pixel feature functors: the HIP/host replacement of the reference's
// user-supplied OpenCL hook `feat_t computeFeature(image, nChannels, width,
// height, coords, treeLeftChildren, treePosteriors, imageNodesID, features,
// featDim)` (test/feature.cl:29-34; contract in src/padenti/padenti.hpp.in:
// 170-175).  A functor is a struct with
//   using ImgType, FeatType, Prep; static constexpr nChannels, FeatDim;
//   static Prep prepare(const ImageView<ImgType,nChannels>&, int x, int y);
//   static FeatType eval_prep(const ImageView<ImgType,nChannels>&, int x, int y,
//                             Prep, const FeatType* params);
//   static FeatType eval(view, x, y, params)
//       == eval_prep(view, x, y, prepare(view, x, y), params);
// compiled for host and device.  `prepare` is the per-pixel part shared by
// every candidate (here the centre depth); the trainer evaluates it once per
// sample and the predictor once per pixel.  Prep must fit in 64 bits.
// Reads outside the image follow the reference sampler (CLK_ADDRESS_CLAMP
// plus an explicit border test in the hook): each functor returns its own
// background value there.
//
// A response goes to the left child when response <= threshold
// (hist_update.cl:154, predict.cl:75).
#pragma once
#include <math.h>
#include <limits>
#include <type_traits>
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
#if defined(__HIP_DEVICE_COMPILE__)
    // 24-bit row multiply (w, y < 2^16) and a 32-bit byte offset from the
    // image base (images are under 4 GiB, checked when they are uploaded):
    // full-rate integer ops and a base + 32-bit offset load.
    const uint32_t i = (__umul24((uint32_t)y, (uint32_t)w) + (uint32_t)x) * (uint32_t)NCH + (uint32_t)c;
    return *(const ImgT *)((const char *)data + i * (uint32_t)sizeof(ImgT));
#else
    return data[((uint32_t)y * (uint32_t)w + (uint32_t)x) * (uint32_t)NCH + (uint32_t)c];
#endif
  }
  PD_HD bool inside(int x, int y) const { return x >= 0 && x < w && y >= 0 && y < h; }
};

// round(q) half away from zero (OpenCL round(), test/feature.cl:50); a
// quotient whose magnitude is not below 2^20 counts as leaving the image,
// which keeps the float->int conversion defined on every target.
PD_HD bool offset_ok(float q) { return fabsf(q) < 1048576.0f; }

// Denominator prepared once per pixel: d and a refined reciprocal y.
struct DivPrep {
  float d, y;
};

PD_HD DivPrep div_prep(float d)
{
#if defined(__HIP_DEVICE_COMPILE__)
  const float y0 = __builtin_amdgcn_rcpf(d);
  const float e = __builtin_fmaf(-d, y0, 1.0f);
  return DivPrep{d, __builtin_fmaf(e, y0, y0)};
#else
  return DivPrep{d, 1.0f / d};
#endif
}

// p / d, IEEE round-to-nearest.  On the device, for d in [2^-16, 2^16] this
// is the compiler's own f32 division sequence for gfx950 (reciprocal, one
// Newton step, quotient, two FMA residual corrections) with its operand
// scaling and special-case fix-up left out, both identities there: the
// quotient bits equal `p / d` for |p| in [2^-60, 2^60], and outside that
// range the rounded offsets computed from it are the same (0, or beyond
// 2^20 / non-finite).  The reciprocal part is shared by all candidates of a
// pixel.  Any other d takes the plain division.
PD_HD float div_by(float p, const DivPrep &dp, bool fast)
{
#if defined(__HIP_DEVICE_COMPILE__)
  if (fast) {
    float q = p * dp.y;
    float r = __builtin_fmaf(-dp.d, q, p);
    q = __builtin_fmaf(r, dp.y, q);
    r = __builtin_fmaf(-dp.d, q, p);
    return __builtin_fmaf(r, dp.y, q);
  }
#endif
  (void)fast;
  return p / dp.d;
}
PD_HD bool div_fast_ok(const DivPrep &dp) { return dp.d >= 0x1p-16f && dp.d <= 0x1p16f; }

// Denominator that is an integer with 1 <= |d| <= 65535 (a uint16 or int16
// depth): d and its correctly rounded reciprocal y = RN(1/d) (an IEEE
// division, once per pixel).  Then one FMA residual correction of p * y is
// already RN(p/d) for every float p whose quotient is a normal number:
// with q0 = RN(p*y), |q0 - p/d| < 2^(e-22) for |p/d| in [2^e, 2^(e+1)), and
// z = q0 + RN(p - d*q0)*y (the residual is exact) satisfies
// z - p/d = (q0 - p/d)(1 - d*y), so |z - p/d| < 2^(e-46); every rounding
// midpoint m near p/d is an odd multiple of 2^(e-25) or 2^(e-24) with a
// 25-bit numerator, so p - m*d is a non-zero multiple of 2^(e-25) and
// |p/d - m| >= 2^(e-25)/|d| > 2^(e-41): no midpoint lies between z and p/d.
// (Subnormal quotients are below 0.5 and round to offset 0 either way.)
struct IntDivPrep {
  float d, y;
};

PD_HD IntDivPrep int_div_prep(float d) { return IntDivPrep{d, 1.0f / d}; }

PD_HD float int_div(float p, const IntDivPrep &dp)
{
#if defined(__HIP_DEVICE_COMPILE__)
  const float q = p * dp.y;
  const float r = __builtin_fmaf(-dp.d, q, p);
  return __builtin_fmaf(r, dp.y, q);
#else
  return p / dp.d;
#endif
}

// Integer offset round(q) (half away from zero) of a quotient, with every
// quotient whose magnitude is not below 2^20 (or NaN) mapped to +-2^20, an
// offset that leaves any image (width, height < 2^16): the same pixel/border
// decision as roundf + offset_ok.  For 0.5 <= |q| <= 2^20, q + copysign(0.5, q)
// rounds to a value with the same integer part as the exact sum (0.5 is a
// multiple of ulp(q) there and a rounding never reaches the next integer),
// so truncation gives round(q); |q| < 0.5 gives 0.
// The device form clamps with one med3 (a quotient of the hooks is never
// NaN: d > 0 and finite parameters) and adds copysign(0.5 - 2^-25, q), which
// needs no |q| < 0.5 case: trunc(q + copysign(0.49999997f, q)) equals the
// host rule for every float32 with |q| <= 2^20 (checked exhaustively).
PD_HD int offset_of(float q)
{
#if defined(__HIP_DEVICE_COMPILE__)
  q = __builtin_amdgcn_fmed3f(q, -1048576.0f, 1048576.0f);
  return (int)(q + copysignf(0.49999997f, q));
#else
  q = fminf(fmaxf(q, -1048576.0f), 1048576.0f);
  const int n = (int)(q + copysignf(0.5f, q));
  return fabsf(q) < 0.5f ? 0 : n;
#endif
}

// Read of channel `ch` at x + round(px/d), y + round(py/d), or `bg` when the
// offset leaves the image.  Only the load is predicated, so lanes whose
// probe leaves the image issue no memory access.
template <typename ImgT, int NCH>
PD_HD float probe_div(const ImageView<ImgT, NCH> &im, int x, int y, const DivPrep &dp, bool fast, float px,
                      float py, int ch, float bg)
{
  const int ox = x + offset_of(div_by(px, dp, fast));
  const int oy = y + offset_of(div_by(py, dp, fast));
  float v = bg;
  if ((uint32_t)ox < (uint32_t)im.w && (uint32_t)oy < (uint32_t)im.h) v = (float)im.at(ox, oy, ch);
  return v;
}

// C1: float depth in metres.  f = I(x + round(u/d(x))) - I(x + round(v/d(x))),
// p = (ux, uy, vx, vy) in pixel*m, out-of-image reads 10 m, d(x) <= 0 -> 0.
struct DepthF32Feature {
  using ImgType = float;
  using FeatType = float;
  using Prep = DivPrep;
  static constexpr int nChannels = 1;
  static constexpr int FeatDim = 4;
  static constexpr int kind = 0;
  PD_HD static Prep prepare(const ImageView<float, 1> &im, int x, int y) { return div_prep(im.at(x, y, 0)); }
  PD_HD static float eval_prep(const ImageView<float, 1> &im, int x, int y, Prep dp, const float *p)
  {
    if (!(dp.d > 0.0f)) return 0.0f;
    const bool fast = div_fast_ok(dp);
    float a = probe_div(im, x, y, dp, fast, p[0], p[1], 0, 10.0f);
    float b = probe_div(im, x, y, dp, fast, p[2], p[3], 0, 10.0f);
    return a - b;
  }
  PD_HD static float eval(const ImageView<float, 1> &im, int x, int y, const float *p)
  {
    return eval_prep(im, x, y, prepare(im, x, y), p);
  }
};

// offset_of(q) for a quotient q = int_div(p, dp) with a positive
// denominator: q then has the sign of the numerator p (zero included), so the
// half added before truncation, copysign(0.5 - 2^-25, p), depends on the
// candidate's parameter only and leaves the per-pixel work (the compiler
// hoists it out of the sample loop).  No clamp: v_cvt_i32_f32 saturates
// (|q| >= 2^31 gives +-(2^31 - 1) or -2^31), and the caller adds the offset
// in unsigned arithmetic, so every |q| >= 2^16 still lands outside any image
// (width, height < 2^16), the decision the host's clamp to +-2^20 makes.  q
// is never NaN (finite parameters, d >= 1).
PD_HD int offset_of_num(float q, float p)
{
#if defined(__HIP_DEVICE_COMPILE__)
  const float t = q + copysignf(0.49999997f, p);
  int r;
  __asm__("v_cvt_i32_f32 %0, %1" : "=v"(r) : "v"(t));
  return r;
#else
  (void)p;
  return offset_of(q);
#endif
}

// Texel of channel `ch` at x + round(px/d), y + round(py/d) with an integer
// depth d > 0 (IntDivPrep) and an integer pixel type, or `bg` when the offset
// leaves the image; the value is returned as an int so that a hook's
// difference of two probes is one integer subtraction (exact, as the float
// one is for values below 2^24).  Only the load is predicated: lanes whose
// probe leaves the image issue no memory access.
template <typename ImgT, int NCH>
PD_HD int probe_int(const ImageView<ImgT, NCH> &im, int x, int y, const IntDivPrep &dp, float px, float py, int ch,
                    int bg)
{
  static_assert(sizeof(ImgT) <= 2 && ImgT(-1) > ImgT(0), "probe_int reads unsigned 8- or 16-bit pixels");
  const uint32_t ox = (uint32_t)x + (uint32_t)offset_of_num(int_div(px, dp), px);
  const uint32_t oy = (uint32_t)y + (uint32_t)offset_of_num(int_div(py, dp), py);
  int v = bg;
  if (ox < (uint32_t)im.w && oy < (uint32_t)im.h) v = (int)im.at((int)ox, (int)oy, ch);
  return v;
}

// C2/C3: uint16 depth in millimetres, float response; p in pixel*mm,
// out-of-image reads 10000 mm, d(x) == 0 -> 0.
struct DepthU16Feature {
  using ImgType = uint16_t;
  using FeatType = float;
  using Prep = IntDivPrep;
  static constexpr int nChannels = 1;
  static constexpr int FeatDim = 4;
  static constexpr int kind = 1;
  static constexpr int kProbeRowA = 1, kProbeRowB = 3;   // row offsets of the two probes
  PD_HD static Prep prepare(const ImageView<uint16_t, 1> &im, int x, int y)
  {
    return int_div_prep((float)im.at(x, y, 0));
  }
  PD_HD static float eval_prep(const ImageView<uint16_t, 1> &im, int x, int y, Prep dp, const float *p)
  {
    if (!(dp.d > 0.0f)) return 0.0f;
    const int a = probe_int(im, x, y, dp, p[0], p[1], 0, 10000);
    const int b = probe_int(im, x, y, dp, p[2], p[3], 0, 10000);
    return (float)(a - b);
  }
  PD_HD static float eval(const ImageView<uint16_t, 1> &im, int x, int y, const float *p)
  {
    return eval_prep(im, x, y, prepare(im, x, y), p);
  }
};

// C4: RGB-D frames as uint16 x 4 (R, G, B, depth mm).  Depth-normalised
// offset channel difference over a channel pair:
// f = I_ca(x + round(u/d)) - I_cb(x + round(v/d)), p = (ux, uy, vx, vy, ca, cb),
// channel = (int)c clamped to [0, 3], out-of-image reads 10000.
struct RGBDFeature {
  using ImgType = uint16_t;
  using FeatType = float;
  using Prep = IntDivPrep;
  static constexpr int nChannels = 4;
  static constexpr int FeatDim = 6;
  static constexpr int kind = 2;
  static constexpr int kProbeRowA = 1, kProbeRowB = 3;
  PD_HD static int chan(float c)
  {
    if (!(c > 0.0f)) return 0;
    if (c >= 3.0f) return 3;
    return (int)c;
  }
  PD_HD static Prep prepare(const ImageView<uint16_t, 4> &im, int x, int y)
  {
    return int_div_prep((float)im.at(x, y, 3));
  }
  PD_HD static float eval_prep(const ImageView<uint16_t, 4> &im, int x, int y, Prep dp, const float *p)
  {
    if (!(dp.d > 0.0f)) return 0.0f;
    const int a = probe_int(im, x, y, dp, p[0], p[1], chan(p[4]), 10000);
    const int b = probe_int(im, x, y, dp, p[2], p[3], chan(p[5]), 10000);
    return (float)(a - b);
  }
  PD_HD static float eval(const ImageView<uint16_t, 4> &im, int x, int y, const float *p)
  {
    return eval_prep(im, x, y, prepare(im, x, y), p);
  }
};

// The reference's own demo hook (test/feature.cl:29-64) with feat_t = short,
// FeatDim 2: depth = (short)I(x); offset = x + (int)round((float)u*500/depth);
// out-of-image -> 10000; response = depth - (d2 ? d2 : 10000) in short.
// depth == 0 (a division by zero in the reference) gives 0.
struct DemoS16Feature {
  using ImgType = uint16_t;
  using FeatType = int16_t;
  using Prep = IntDivPrep;   // d = (float)(short)I(x), exact, |d| <= 32768
  static constexpr int nChannels = 1;
  static constexpr int FeatDim = 2;
  static constexpr int kind = 3;
  static constexpr int kProbeRowA = 1, kProbeRowB = 1;   // one probe, row offset p[1]
  PD_HD static Prep prepare(const ImageView<uint16_t, 1> &im, int x, int y)
  {
    return int_div_prep((float)(int16_t)im.at(x, y, 0));
  }
  PD_HD static int16_t eval_prep(const ImageView<uint16_t, 1> &im, int x, int y, Prep dp, const int16_t *p)
  {
    const int16_t depth = (int16_t)dp.d;
    if (depth == 0) return 0;
    const int ox = x + offset_of(int_div((float)p[0] * 500.0f, dp));
    const int oy = y + offset_of(int_div((float)p[1] * 500.0f, dp));
    int16_t d2 = (int16_t)10000;
    if ((uint32_t)ox < (uint32_t)im.w && (uint32_t)oy < (uint32_t)im.h) d2 = (int16_t)im.at(ox, oy, 0);
    return (int16_t)(depth - (d2 ? d2 : (int16_t)10000));
  }
  PD_HD static int16_t eval(const ImageView<uint16_t, 1> &im, int x, int y, const int16_t *p)
  {
    return eval_prep(im, x, y, prepare(im, x, y), p);
  }
};

// Prep <-> the two 32-bit words (z, w) of the 16-byte sample record.
template <typename P>
PD_HD void pack_prep(P v, uint32_t &z, uint32_t &w)
{
  static_assert(sizeof(P) <= 8, "Prep must fit in 64 bits");
  uint32_t t[2] = {0u, 0u};
  memcpy(t, &v, sizeof(P));
  z = t[0];
  w = t[1];
}
template <typename P>
PD_HD P unpack_prep(uint32_t z, uint32_t w)
{
  uint32_t t[2] = {z, w};
  P v;
  memcpy(&v, t, sizeof(P));
  return v;
}

// ---------------------------------------------------------------------------
// Auto-context hooks (the reference passes treeLeftChildren, treePosteriors
// and the image of node ids to every hook call, hist_update.cl:122-124,
// predict.cl:70-74).  A functor opts in with
//   static constexpr bool kAutoContext = true;
//   static FeatType eval_ctx(view, x, y, Prep, const FeatType *params,
//                            const AutoContext &ctx);
// and may then read the tree being trained (or predicted with) and the node
// id of any pixel of its image at the current depth (-1 outside the image
// and for pixels that are not routed: unlabelled in training, masked out in
// prediction).
// ---------------------------------------------------------------------------
struct AutoContext {
  const int32_t *left;      // left children of the tree (breadth-first heap)
  const float *post;        // posteriors [node][n_classes]
  uint32_t n_classes;
  const int32_t *node_ids;  // this image's node id per pixel, row-major
  int w, h;
  PD_HD int32_t node_at(int x, int y) const
  {
    return (x < 0 || x >= w || y < 0 || y >= h) ? -1 : node_ids[(size_t)y * (size_t)w + (size_t)x];
  }
  PD_HD float posterior(int32_t node, uint32_t c) const
  {
    return (node < 0 || c >= n_classes) ? 0.0f : post[(size_t)node * n_classes + c];
  }
};

// Kernel-side view over all images: node maps packed back to back, image i's
// map at map_off[i].  All pointers are null for hooks without context.
struct AcArgs {
  const int32_t *left;
  const float *post;
  uint32_t n_classes;
  const int32_t *maps;
  const uint64_t *map_off;
};

template <class F, class = void>
struct uses_context : std::false_type {};
template <class F>
struct uses_context<F, std::void_t<decltype(F::kAutoContext)>> : std::integral_constant<bool, F::kAutoContext> {};

// The hook call of every kernel: eval_prep, or eval_ctx with image `img`'s context.
template <class F>
PD_HD typename F::FeatType eval_hook(const ImageView<typename F::ImgType, F::nChannels> &view, int x, int y,
                                     typename F::Prep prep, const typename F::FeatType *p, const AcArgs &ac,
                                     uint32_t img)
{
  if constexpr (uses_context<F>::value) {
    const AutoContext ctx{ac.left, ac.post, ac.n_classes, ac.maps + ac.map_off[img], view.w, view.h};
    return F::eval_ctx(view, x, y, prep, p, ctx);
  } else {
    (void)ac;
    (void)img;
    return F::eval_prep(view, x, y, prep, p);
  }
}

// Threshold padding value: compares "left" against every response, so a
// padded slot never counts as "right" for a finite response (+inf for the
// floating types, the largest value for the integer ones).
template <typename T>
PD_HD T pad_threshold()
{
  if constexpr (std::is_floating_point<T>::value) return (T)__builtin_inf();
  else return std::numeric_limits<T>::max();
}

}  // namespace pd
