// TrainingSetImage: image + labels (0 = unlabelled, 1..nClasses) + sampled
// pixel indices; per-image priors over the labelled samples
// (src/padenti/training_set_image_impl.hpp:24-60).
#pragma once
#include <vector>
#include "image.hpp"

template <typename type, unsigned int nChannels>
class TrainingSetImage : public Image<type, nChannels>
{
protected:
  std::vector<unsigned int> m_samples;
  std::vector<unsigned char> m_labels;
  unsigned int m_nClasses;
  std::vector<float> m_priors;

public:
  TrainingSetImage(const type *data, unsigned int width, unsigned int height, const unsigned char *labels,
                   unsigned int nClasses, const unsigned int *samples, unsigned int nSamples)
    : Image<type, nChannels>(data, width, height),
      m_samples(samples, samples + nSamples),
      m_labels(labels, labels + (size_t)width * height),
      m_nClasses(nClasses),
      m_priors(nClasses, 0.0f)
  {
    std::vector<unsigned int> cnt(nClasses, 0u);
    unsigned int tot = 0;
    for (unsigned int i = 0; i < nSamples; i++) {
      const unsigned char l = labels[samples[i]];
      if (l) {
        cnt[l - 1]++;
        tot++;
      }
    }
    for (unsigned int c = 0; c < nClasses; c++) m_priors[c] = (float)cnt[c] / (float)tot;
  }
  const unsigned char *getLabels() const { return m_labels.data(); }
  const unsigned int *getSamples() const { return m_samples.data(); }
  unsigned int getNSamples() const { return (unsigned int)m_samples.size(); }
  const float *getPriors() const { return m_priors.data(); }
};
