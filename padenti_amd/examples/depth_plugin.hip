// Example of a user-supplied feature hook compiled as a plugin.  It restates
// the float-depth comparison of the small configuration in a user's own
// code (f = I(x + round(u/d)) - I(x + round(v/d)), 10 m outside, d <= 0 -> 0)
// so that trees trained through it can be checked against the CPU oracle.
#include <padenti/pd_plugin.hpp>

struct UserDepthFeature {
  using ImgType = float;
  using FeatType = float;
  using Prep = pd::DivPrep;
  static constexpr int nChannels = 1;
  static constexpr int FeatDim = 4;
  PD_HD static Prep prepare(const pd::ImageView<float, 1> &im, int x, int y) { return pd::div_prep(im.at(x, y, 0)); }
  PD_HD static float read(const pd::ImageView<float, 1> &im, int x, int y, const Prep &dp, bool fast, float px,
                          float py)
  {
    return pd::probe_div(im, x, y, dp, fast, px, py, 0, 10.0f);
  }
  PD_HD static float eval_prep(const pd::ImageView<float, 1> &im, int x, int y, Prep dp, const float *p)
  {
    if (!(dp.d > 0.0f)) return 0.0f;
    const bool fast = pd::div_fast_ok(dp);
    return read(im, x, y, dp, fast, p[0], p[1]) - read(im, x, y, dp, fast, p[2], p[3]);
  }
  PD_HD static float eval(const pd::ImageView<float, 1> &im, int x, int y, const float *p)
  {
    return eval_prep(im, x, y, prepare(im, x, y), p);
  }
};

PD_DEFINE_FEATURE(user_depth, UserDepthFeature)
