// An auto-context feature hook (the reference's hook receives the tree
// being trained and the image of node ids, hist_update.cl:122-124): the
// depth difference to an offset pixel plus a weighted posterior of that
// pixel's current node,
//   f = (I(x + ox, y + oy) - I(x, y)) + w * P(c | node of (x + ox, y + oy))
// with params (ox, oy, c, w): integer offsets (truncated), c = class (mod
// n_classes), w a weight; outside the image I = 10 and the posterior is 0.
#include <padenti/pd_plugin.hpp>

struct DepthContextFeature {
  using ImgType = float;
  using FeatType = float;
  using Prep = uint32_t;
  static constexpr int nChannels = 1;
  static constexpr int FeatDim = 4;
  static constexpr bool kAutoContext = true;
  PD_HD static Prep prepare(const pd::ImageView<float, 1> &, int, int) { return 0u; }
  PD_HD static float eval_ctx(const pd::ImageView<float, 1> &im, int x, int y, Prep, const float *p,
                              const pd::AutoContext &ctx)
  {
    const int ox = x + (int)p[0], oy = y + (int)p[1];
    const float a = im.inside(ox, oy) ? im.at(ox, oy, 0) : 10.0f;
    const uint32_t c = (uint32_t)p[2] % ctx.n_classes;
    const float q = ctx.posterior(ctx.node_at(ox, oy), c);
    const float d = a - im.at(x, y, 0);
    const float wq = p[3] * q;
    return d + wq;
  }
};

PD_DEFINE_FEATURE(depth_context, DepthContextFeature)
