// HIPTreeTrainer: the drop-in for CLTreeTrainer(featureKernelPath, useCPU)
// (src/padenti/cl_tree_trainer.hpp:115, train cl_tree_trainer_impl.hpp:203-255)
// over the C ABI of libpadenti_hip.so.  The feature hook is a compiled HIP
// functor: one of the built-in kinds (PD_FEAT_*), or a plugin's create
// function (PD_DEFINE_FEATURE in pd_plugin.hpp).  Errors throw
// std::runtime_error with the library's message.
#pragma once
#include <stdexcept>
#include <string>
#include <vector>

#include "../padenti_hip.h"
#include "tree_trainer.hpp"

typedef int (*pd_trainer_factory)(pd_trainer **out, int device, uint32_t n_classes);

namespace pd_cxx {
inline void check(int rc)
{
  if (rc < 0) throw std::runtime_error(std::string("padenti_hip: ") + pd_last_error());
}
}  // namespace pd_cxx

template <typename ImgType, unsigned int nChannels, typename FeatType, unsigned int FeatDim, unsigned int nClasses>
class HIPTreeTrainer : public TreeTrainer<ImgType, nChannels, FeatType, FeatDim, nClasses>
{
private:
  pd_trainer *m_tr;
  const void *m_uploaded;
  size_t m_uploadedImages;
  std::vector<pd_level_stats> m_stats;

  HIPTreeTrainer(const HIPTreeTrainer &);
  HIPTreeTrainer &operator=(const HIPTreeTrainer &);

  void upload(const TrainingSet<ImgType, nChannels> &ts)
  {
    const std::vector<TrainingSetImage<ImgType, nChannels> > &imgs = ts.getImages();
    const uint32_t n = (uint32_t)imgs.size();
    std::vector<uint64_t> img_off(n), lab_off(n), smp_off(n);
    std::vector<uint32_t> w(n), h(n), ns(n);
    uint64_t pix = 0, lab = 0, smp = 0;
    for (uint32_t i = 0; i < n; i++) {
      w[i] = imgs[i].getWidth();
      h[i] = imgs[i].getHeight();
      ns[i] = imgs[i].getNSamples();
      img_off[i] = pix;
      lab_off[i] = lab;
      smp_off[i] = smp;
      pix += (uint64_t)w[i] * h[i] * nChannels;
      lab += (uint64_t)w[i] * h[i];
      smp += ns[i];
    }
    std::vector<ImgType> pixels(pix);
    std::vector<uint8_t> labels(lab);
    std::vector<uint32_t> samples(smp);
    for (uint32_t i = 0; i < n; i++) {
      std::copy(imgs[i].getData(), imgs[i].getData() + (size_t)w[i] * h[i] * nChannels, &pixels[img_off[i]]);
      std::copy(imgs[i].getLabels(), imgs[i].getLabels() + (size_t)w[i] * h[i], &labels[lab_off[i]]);
      std::copy(imgs[i].getSamples(), imgs[i].getSamples() + ns[i], samples.data() + smp_off[i]);
    }
    pd_cxx::check(pd_trainer_set_data(m_tr, n, pixels.data(), img_off.data(), w.data(), h.data(), labels.data(),
                                      lab_off.data(), samples.data(), smp_off.data(), ns.data()));
    m_uploaded = &ts;
    m_uploadedImages = n;
  }

public:
  explicit HIPTreeTrainer(int featureKind, int device = 0) : m_tr(NULL), m_uploaded(NULL), m_uploadedImages(0)
  {
    pd_cxx::check(pd_trainer_create(&m_tr, device, featureKind, nClasses));
  }
  explicit HIPTreeTrainer(pd_trainer_factory create, int device = 0)
    : m_tr(NULL), m_uploaded(NULL), m_uploadedImages(0)
  {
    pd_cxx::check(create(&m_tr, device, nClasses));
  }
  ~HIPTreeTrainer()
  {
    if (m_tr) pd_trainer_destroy(m_tr);
  }

  // Device histogram budget (the reference's GLOBAL_HISTOGRAM_MAX_SIZE).
  void setHistogramBudget(uint64_t bytes) { pd_cxx::check(pd_trainer_set_hist_budget(m_tr, bytes)); }
  // Per-level timings of the last train() call.
  const std::vector<pd_level_stats> &levelStats() const { return m_stats; }

  void train(Tree<FeatType, FeatDim, nClasses> &tree, const TrainingSet<ImgType, nChannels> &trainingSet,
             const TreeTrainerParameters<FeatType, FeatDim> &params, unsigned int startDepth, unsigned int endDepth)
  {
    if (startDepth != 1) throw std::runtime_error("Starting depth must be equal to 1");
    if (m_uploaded != &trainingSet || m_uploadedImages != trainingSet.getImages().size()) upload(trainingSet);
    pd_train_params p = {};
    p.n_features = params.nFeatures;
    p.n_thresholds = params.nThresholds;
    for (unsigned int j = 0; j < FeatDim; j++) {
      p.feat_low[j] = (double)params.featLowBounds[j];
      p.feat_up[j] = (double)params.featUpBounds[j];
    }
    p.thr_low = (double)params.thrLowBound;
    p.thr_up = (double)params.thrUpBound;
    p.per_leaf_samples_thr = params.perLeafSamplesThr;
    m_stats.assign(endDepth > 1 ? endDepth - 1 : 1, pd_level_stats());
    pd_cxx::check(pd_trainer_train(m_tr, &p, tree.getID(), tree.getDepth(), endDepth, PD_MODE_LEVEL_BATCHED,
                                   trainingSet.getPriors(), tree.getLeftChildren(), tree.getFeatures(),
                                   tree.getThresholds(), tree.getPosteriors(), NULL, NULL, NULL, m_stats.data()));
  }
};
