// The C++ drop-in API used the way the reference's demo uses CLTreeTrainer /
// CLClassifier (test/test_tree_trainer.cpp:53-98, test/test_classifier.cpp):
// build a training set, train trees with a built-in hook and with a plugin
// hook, save / load them as XML, classify a held-out image; build a training
// set from a directory of PNGs (test/test_training_set.cpp:34-39) with both
// samplers.  Writes the
// arrays it produced to <outdir> for the test-suite to compare with the CPU
// oracle.   usage: api_demo <outdir>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include <sys/stat.h>

#include <padenti/padenti.hpp>

extern "C" int pd_synth_image(int kind, uint64_t seed, uint32_t img, uint32_t w, uint32_t h, uint32_t n_classes,
                              void *pixels, uint8_t *labels);
extern "C" int pd_synth_samples(uint64_t seed, uint32_t img, const uint8_t *labels, uint32_t w, uint32_t h,
                                uint32_t n, uint32_t *out);
extern "C" int pd_trainer_create_user_depth(pd_trainer **out, int device, uint32_t n_classes);
extern "C" int pd_forest_create_user_depth(pd_forest **out, int device, uint32_t n_classes);

typedef TrainingSet<float, 1> TrainingSetT;
typedef Tree<float, 4, 4> TreeT;
typedef TreeTrainerParameters<float, 4> ParamsT;
typedef HIPTreeTrainer<float, 1, float, 4, 4> TrainerT;
typedef HIPClassifier<float, 1, float, 4, 4> ClassifierT;

static void dump(const std::string &path, const void *p, size_t bytes)
{
  std::ofstream f(path.c_str(), std::ios::binary);
  f.write((const char *)p, (std::streamsize)bytes);
}

static bool same_tree(const TreeT &a, const TreeT &b)
{
  const size_t n = a.getNNodes();
  return n == b.getNNodes() && !memcmp(a.getLeftChildren(), b.getLeftChildren(), n * 4) &&
         !memcmp(a.getFeatures(), b.getFeatures(), n * 16) && !memcmp(a.getThresholds(), b.getThresholds(), n * 4) &&
         !memcmp(a.getPosteriors(), b.getPosteriors(), n * 16);
}

int main(int argc, char **argv)
{
  if (argc != 2) {
    fprintf(stderr, "usage: %s <outdir>\n", argv[0]);
    return 2;
  }
  const std::string out = argv[1];
  const unsigned W = 160, H = 120, NS = 128, NIMG = 8;
  TrainingSetT ts(4);
  std::vector<float> px(W * H);
  std::vector<uint8_t> lab(W * H);
  std::vector<uint32_t> smp(NS);
  for (unsigned i = 0; i < NIMG; i++) {
    pd_synth_image(0, 0, i, W, H, 4, px.data(), lab.data());
    pd_synth_samples(0, i, lab.data(), W, H, NS, smp.data());
    ts << TrainingSetImage<float, 1>(px.data(), W, H, lab.data(), 4, smp.data(), NS);
  }
  ParamsT params;
  params.nFeatures = 64;
  params.nThresholds = 10;
  params.computeFRange = false;
  for (int j = 0; j < 4; j++) {
    params.featLowBounds[j] = -60.0f;
    params.featUpBounds[j] = 60.0f;
  }
  params.thrLowBound = -0.5f;
  params.thrUpBound = 0.5f;
  params.randomThrSampling = true;
  params.perLeafSamplesThr = 10;

  try {
    TreeT tree(0, 6), tree_plugin(0, 6);
    TrainerT trainer(PD_FEAT_DEPTH_F32);
    trainer.train(tree, ts, params, 1, 6);
    TrainerT trainer_plugin(&pd_trainer_create_user_depth);
    trainer_plugin.train(tree_plugin, ts, params, 1, 6);
    if (!same_tree(tree, tree_plugin)) {
      fprintf(stderr, "plugin hook trained a different tree\n");
      return 1;
    }
    bool threw = false;
    try {
      trainer.train(tree, ts, params, 2, 6);
    } catch (const std::runtime_error &) {
      threw = true;
    }
    if (!threw) return 1;

    // launch tuning changes timing only
    {
      TreeT tuned(0, 6);
      TrainerT t2(PD_FEAT_DEPTH_F32);
      t2.setTuning(16, 64);
      t2.train(tuned, ts, params, 1, 6);
      if (!same_tree(tree, tuned)) {
        fprintf(stderr, "tuned launch trained a different tree\n");
        return 1;
      }
      // 16-bit bins, no candidate order, small chunks: the same tree again
      TreeT layout(0, 6);
      TrainerT t3(PD_FEAT_DEPTH_F32);
      t3.setWideBins(false);
      t3.setProbeOrder(false);
      t3.setChunk(128);
      t3.train(layout, ts, params, 1, 6);
      if (!same_tree(tree, layout)) {
        fprintf(stderr, "histogram layout knobs trained a different tree\n");
        return 1;
      }
    }
    // learned split proposal: records from a plain training, the network
    // trained on them, and a proposal that keeps every candidate (= plain)
    {
      const uint64_t cap = 1u << 16;
      void *rows = NULL, *gain = NULL, *rel = NULL;
      pd_cxx::check(pd_dev_alloc(0, cap * 64, &rows));
      pd_cxx::check(pd_dev_alloc(0, cap * 4, &gain));
      pd_cxx::check(pd_dev_alloc(0, cap * 4, &rel));
      TrainerT t3(PD_FEAT_DEPTH_F32);
      t3.setRecordSink((uint16_t *)rows, (float *)gain, (float *)rel, cap, 100);
      TreeT plain(0, 6);
      t3.train(plain, ts, params, 1, 6);
      const uint64_t n = t3.recordsWritten();
      t3.setRecordSink(NULL, NULL, NULL, 0, 0);
      SplitProposalNet net(3);
      const std::vector<float> loss = net.train((const uint16_t *)rows, (const float *)rel, n, 1024, 2);
      if (n < 1000 || !(loss[1] == loss[1]) || !same_tree(tree, plain)) {
        fprintf(stderr, "record sink / network training failed (%llu records)\n", (unsigned long long)n);
        return 1;
      }
      t3.setProposal(net, params.nFeatures, 0, true);
      t3.setMode(PD_MODE_PROPOSAL);
      TreeT prop(0, 6);
      t3.train(prop, ts, params, 1, 6);
      if (!same_tree(tree, prop)) {
        fprintf(stderr, "proposal keeping every candidate trained a different tree\n");
        return 1;
      }
      pd_dev_free(rows);
      pd_dev_free(gain);
      pd_dev_free(rel);
    }
    tree.save(out + "/tree.xml", -1, true);
    tree.save(out + "/tree_default.xml");
    TreeT loaded;
    loaded.load(out + "/tree.xml");
    if (!same_tree(tree, loaded)) {
      fprintf(stderr, "lossless XML round trip changed the tree\n");
      return 1;
    }
    const size_t n = tree.getNNodes();
    dump(out + "/left.bin", tree.getLeftChildren(), n * 4);
    dump(out + "/features.bin", tree.getFeatures(), n * 16);
    dump(out + "/thresholds.bin", tree.getThresholds(), n * 4);
    dump(out + "/posteriors.bin", tree.getPosteriors(), n * 16);
    dump(out + "/priors.bin", ts.getPriors(), 16);

    pd_synth_image(0, 0, 1000000, W, H, 4, px.data(), lab.data());
    Image<float, 1> img(px.data(), W, H);
    ClassifierT clf(PD_FEAT_DEPTH_F32), clf_plugin(&pd_forest_create_user_depth);
    clf << tree;
    clf_plugin << tree_plugin;
    Image<float, 4> post(W, H), post_plugin(W, H);
    Image<int, 1> leaves(W, H);
    clf.predict(img, post);
    clf_plugin.predict(img, post_plugin);
    clf.predict(0, img, leaves);
    if (memcmp(post.getData(), post_plugin.getData(), (size_t)W * H * 4 * 4)) {
      fprintf(stderr, "plugin hook predicted differently\n");
      return 1;
    }
    Image<unsigned char, 1> mask(W, H);
    for (unsigned i = 0; i < W * H; i++) mask.getData()[i] = (unsigned char)(i % 3 != 0);
    Image<float, 4> post_masked(W, H);
    clf.predict(img, post_masked, mask);
    dump(out + "/image.bin", px.data(), (size_t)W * H * 4);
    dump(out + "/post.bin", post.getData(), (size_t)W * H * 4 * 4);
    dump(out + "/post_masked.bin", post_masked.getData(), (size_t)W * H * 4 * 4);
    dump(out + "/leaves.bin", leaves.getData(), (size_t)W * H * 4);
  } catch (const std::exception &e) {
    fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  try {  // a training set from a directory of depth / RGB-label PNGs
    static const unsigned char RGB2LABEL[8][3] = {{255, 0, 0}, {0, 0, 255}, {0, 255, 0}, {255, 255, 0},
                                                  {255, 0, 255}, {0, 255, 255}, {128, 64, 32}, {250, 250, 250}};
    const std::string dir = out + "/ds";
    mkdir(dir.c_str(), 0755);
    const unsigned DW = 96, DH = 64, ND = 5;
    std::vector<unsigned short> depth(DW * DH);
    std::vector<uint8_t> dl(DW * DH), bgr(DW * DH * 3);
    for (unsigned i = 0; i < ND; i++) {
      pd_synth_image(1, 0, i, DW, DH, 8, depth.data(), dl.data());
      for (unsigned k = 0; k < DW * DH; k++)
        for (unsigned c = 0; c < 3; c++) bgr[3 * k + c] = dl[k] ? RGB2LABEL[dl[k] - 1][2 - c] : 0;
      char name[64];
      snprintf(name, sizeof(name), "/s%02u_depth.png", i);
      pd_cxx::io_check(pd_png_write((dir + name).c_str(), DW, DH, 1, 16, depth.data()));
      snprintf(name, sizeof(name), "/s%02u_labels.png", i);
      pd_cxx::io_check(pd_png_write((dir + name).c_str(), DW, DH, 3, 8, bgr.data()));
    }
    PNGImageLoader<unsigned short, 1> depthLoader;
    RGBLabelsLoader labelsLoader(RGB2LABEL, 8);
    UniformImageSampler<unsigned short, 1> uniform(300, 7);
    ClassWeightedImageSampler<unsigned short, 1> weighted(300, 8, 7);
    TrainingSet<unsigned short, 1> tsu(dir, "_depth.png", "_labels.png", 8, depthLoader, labelsLoader, uniform);
    TrainingSet<unsigned short, 1> tsw(dir, "_depth.png", "_labels.png", 8, depthLoader, labelsLoader, weighted);
    if (tsu.getNImages() != ND || tsw.getNImages() != ND) {
      fprintf(stderr, "directory training set has %u / %u images\n", tsu.getNImages(), tsw.getNImages());
      return 1;
    }
    std::vector<uint32_t> counts, all;
    for (unsigned i = 0; i < ND; i++)
      for (int k = 0; k < 2; k++) {
        const TrainingSetImage<unsigned short, 1> &im = (k ? tsw : tsu).getImages()[i];
        counts.push_back(im.getNSamples());
        all.insert(all.end(), im.getSamples(), im.getSamples() + im.getNSamples());
      }
    dump(out + "/ds_counts.bin", counts.data(), counts.size() * 4);
    dump(out + "/ds_samples.bin", all.data(), all.size() * 4);
    dump(out + "/ds_priors.bin", tsu.getPriors(), 8 * 4);
  } catch (const std::exception &e) {
    fprintf(stderr, "dataset error: %s\n", e.what());
    return 1;
  }
  printf("api_demo ok\n");
  return 0;
}
