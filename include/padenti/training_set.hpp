// TrainingSet: ordered images and the running mean of their priors
// (src/padenti/training_set_impl.hpp:32-39, 157-170).  The directory-scan
// constructor (OpenCV loaders + Boost samplers) is not provided: images and
// their samples enter through operator<<.
#pragma once
#include <vector>
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
