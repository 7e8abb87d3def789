// TrainingSet: ordered images and the running mean of their priors
// (src/padenti/training_set_impl.hpp:32-39, 157-170), and the directory
// constructor (training_set_impl.hpp:43-133): every regular file of tsPath
// ending with dataSuffix whose <prefix><labelsSuffix> exists is a pair, taken
// in file-name order; pairs of different sizes are skipped with a warning;
// all images are sampled in one batch; the priors are the FP64 mean of the
// per-image priors cast to float.
#pragma once
#include <cstdio>
#include <string>
#include <vector>
#include "image_loader.hpp"
#include "image_sampler.hpp"
#include "training_set_image.hpp"

template <typename type, unsigned int nChannels>
class TrainingSet
{
private:
  unsigned int m_nImages;
  unsigned int m_nClasses;
  std::vector<float> m_priors;
  std::vector<TrainingSetImage<type, nChannels> > m_images;

public:
  explicit TrainingSet(unsigned int nClasses) : m_nImages(0), m_nClasses(nClasses), m_priors(nClasses, 0.0f) {}
  TrainingSet(const std::string &tsPath, const std::string &dataSuffix, const std::string &labelsSuffix,
              unsigned int nClasses, ImageLoader<type, nChannels> &dataLoader,
              ImageLoader<unsigned char, 1> &labelsLoader, const ImageSampler<type, nChannels> &sampler)
    : m_nImages(0), m_nClasses(nClasses), m_priors(nClasses, 0.0f)
  {
    uint32_t n_pairs = 0;
    uint64_t need = 0;
    pd_cxx::io_check(pd_list_pairs(tsPath.c_str(), dataSuffix.c_str(), labelsSuffix.c_str(), NULL, 0, &n_pairs, &need));
    std::vector<char> buf(need + 1);
    pd_cxx::io_check(
      pd_list_pairs(tsPath.c_str(), dataSuffix.c_str(), labelsSuffix.c_str(), buf.data(), need, &n_pairs, &need));
    std::vector<Image<type, nChannels> > imgs;
    std::vector<Image<unsigned char, 1> > labs;
    const char *p = buf.data();
    for (uint32_t i = 0; i < n_pairs; i++) {
      const std::string dataPath(p);
      p += dataPath.size() + 1;
      const std::string labelsPath(p);
      p += labelsPath.size() + 1;
      Image<type, nChannels> img = dataLoader.load(dataPath);
      Image<unsigned char, 1> lab = labelsLoader.load(labelsPath);
      if (img.getWidth() != lab.getWidth() || img.getHeight() != lab.getHeight()) {
        std::fprintf(stderr, "warning: skip image pair %s - %s due to different size (%uX%u, %uX%u)\n",
                     dataPath.c_str(), labelsPath.c_str(), img.getWidth(), img.getHeight(), lab.getWidth(),
                     lab.getHeight());
        continue;
      }
      imgs.push_back(img);
      labs.push_back(lab);
    }
    std::vector<const type *> dp;
    std::vector<const unsigned char *> lp;
    std::vector<unsigned int> ws, hs;
    for (size_t i = 0; i < imgs.size(); i++) {
      dp.push_back(imgs[i].getData());
      lp.push_back(labs[i].getData());
      ws.push_back(imgs[i].getWidth());
      hs.push_back(imgs[i].getHeight());
    }
    std::vector<std::vector<unsigned int> > smp;
    sampler.sampleBatch(imgs.size(), dp.data(), lp.data(), ws.data(), hs.data(), smp);
    std::vector<double> acc(nClasses, 0.0);
    for (size_t i = 0; i < imgs.size(); i++) {
      m_images.push_back(TrainingSetImage<type, nChannels>(dp[i], ws[i], hs[i], lp[i], nClasses, smp[i].data(),
                                                           (unsigned int)smp[i].size()));
      for (unsigned int c = 0; c < nClasses; c++) acc[c] += m_images.back().getPriors()[c];
    }
    m_nImages = (unsigned int)m_images.size();
    for (unsigned int c = 0; c < nClasses; c++) m_priors[c] = static_cast<float>(acc[c] / m_nImages);
  }
  const std::vector<TrainingSetImage<type, nChannels> > &getImages() const { return m_images; }
  TrainingSet &operator<<(const TrainingSetImage<type, nChannels> &image)
  {
    m_images.push_back(image);
    for (unsigned int c = 0; c < m_nClasses; c++) {
      const double pC = m_priors[c], pI = image.getPriors()[c];
      m_priors[c] = static_cast<float>(pC * m_nImages / (m_nImages + 1) + pI / (m_nImages + 1));
    }
    m_nImages++;
    return *this;
  }
  unsigned int getNImages() const { return m_nImages; }
  unsigned int getNClasses() const { return m_nClasses; }
  const float *getPriors() const { return m_priors.data(); }
};
