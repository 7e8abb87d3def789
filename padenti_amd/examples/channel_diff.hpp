// A feature hook written once for any image element type, channel count and
// FeatType (the reference's plugin contract over every FeatType of
// src/padenti/cl_feat_fmt_traits.hpp:38-109 and images of more than four
// channels, cl_tree_trainer_impl.hpp:134-144):
//   params (ox, oy, c1, c2):  a = I(x + (int)ox, y + (int)oy, c1 mod nCh),
//                             0 outside the image;  b = I(x, y, c2 mod nCh);
//   response = (FeatType)((FeatType)a - (FeatType)b)
// in C/C++ conversion rules (modular for unsigned targets; the channel
// parameters must be non-negative).  Compiled with the light kernel set.
#pragma once
#define PD_KERNEL_SET_LIGHT 1
#include <padenti/pd_plugin.hpp>

template <typename ImgT, int NCH, typename FeatT>
struct ChannelDiff {
  using ImgType = ImgT;
  using FeatType = FeatT;
  using Prep = uint32_t;
  static constexpr int nChannels = NCH;
  static constexpr int FeatDim = 4;
  PD_HD static Prep prepare(const pd::ImageView<ImgT, NCH> &, int, int) { return 0u; }
  PD_HD static FeatT eval_prep(const pd::ImageView<ImgT, NCH> &im, int x, int y, Prep, const FeatT *p)
  {
    const int ox = x + (int)p[0], oy = y + (int)p[1];
    const int c1 = (int)((uint32_t)p[2] % (uint32_t)NCH), c2 = (int)((uint32_t)p[3] % (uint32_t)NCH);
    const FeatT a = im.inside(ox, oy) ? (FeatT)im.at(ox, oy, c1) : (FeatT)0;
    const FeatT b = (FeatT)im.at(x, y, c2);
    return (FeatT)(a - b);
  }
  PD_HD static FeatT eval(const pd::ImageView<ImgT, NCH> &im, int x, int y, const FeatT *p)
  {
    return eval_prep(im, x, y, prepare(im, x, y), p);
  }
};
