// Pixel samplers: the ImageSampler interface (src/padenti/image_sampler.hpp:
// 27-62) and the reference's two samplers, computed on the GPU with the
// reference's draws (MT19937 seeded with crc32(image bytes) ^ seed):
//   UniformImageSampler         uniform_image_sampler_impl.hpp:27-63
//   ClassWeightedImageSampler   classweighted_image_sampler_impl.hpp:26-123
// sampleBatch() samples many images in one device call (the directory-built
// TrainingSet uses it); its default, for user samplers, calls sample() per
// image.  An image without labelled pixels gets no sample.
#pragma once
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../padenti_hip.h"

template <typename type, unsigned int nChannels>
class ImageSampler
{
protected:
  unsigned int m_nSamples;
  unsigned int m_seed;

public:
  ImageSampler(unsigned int nSamples, unsigned int seed = 0) : m_nSamples(nSamples), m_seed(seed) {}
  virtual ~ImageSampler() {}
  // samples must hold capacity() values; returns the number written
  virtual unsigned int sample(const type *data, const unsigned char *labels, unsigned int width,
                              unsigned int height, unsigned int *samples) const = 0;
  virtual unsigned int capacity() const { return m_nSamples; }
  virtual void sampleBatch(size_t n, const type *const *data, const unsigned char *const *labels,
                           const unsigned int *widths, const unsigned int *heights,
                           std::vector<std::vector<unsigned int> > &out) const
  {
    out.assign(n, std::vector<unsigned int>());
    std::vector<unsigned int> buf(capacity() ? capacity() : 1);
    for (size_t i = 0; i < n; i++) {
      const unsigned int k = sample(data[i], labels[i], widths[i], heights[i], buf.data());
      out[i].assign(buf.begin(), buf.begin() + k);
    }
  }
};

namespace pd_cxx {
// pd_sample_images over host images; out[i] = samples of image i
template <typename type, unsigned int nChannels>
void sample_on_device(int device, const pd_sampler &s, size_t n, const type *const *data,
                      const unsigned char *const *labels, const unsigned int *widths, const unsigned int *heights,
                      std::vector<std::vector<unsigned int> > &out)
{
  out.assign(n, std::vector<unsigned int>());
  if (n == 0) return;
  uint32_t cap = 0;
  if (pd_sampler_capacity(&s, &cap) < 0) throw std::runtime_error(std::string("padenti_hip: ") + pd_last_error());
  std::vector<uint64_t> doff(n), dbytes(n), loff(n);
  uint64_t dt = 0, lt = 0;
  for (size_t i = 0; i < n; i++) {
    const uint64_t px = (uint64_t)widths[i] * heights[i];
    doff[i] = dt;
    dbytes[i] = px * nChannels * sizeof(type);
    dt += dbytes[i];
    loff[i] = lt;
    lt += px;
  }
  std::vector<unsigned char> d(dt ? dt : 1), l(lt ? lt : 1);
  for (size_t i = 0; i < n; i++) {
    std::memcpy(&d[doff[i]], data[i], dbytes[i]);
    std::memcpy(&l[loff[i]], labels[i], (size_t)widths[i] * heights[i]);
  }
  std::vector<uint32_t> res((size_t)n * cap + 1), n_out(n);
  if (pd_sample_images(device, &s, (uint32_t)n, d.data(), doff.data(), dbytes.data(), l.data(), loff.data(), widths,
                       heights, res.data(), n_out.data(), NULL) < 0)
    throw std::runtime_error(std::string("padenti_hip: ") + pd_last_error());
  for (size_t i = 0; i < n; i++) out[i].assign(res.begin() + (size_t)i * cap, res.begin() + (size_t)i * cap + n_out[i]);
}
}  // namespace pd_cxx

template <typename type, unsigned int nChannels>
class GPUImageSampler : public ImageSampler<type, nChannels>
{
protected:
  pd_sampler m_desc;
  std::vector<float> m_weights;
  int m_device;

  GPUImageSampler(unsigned int kind, unsigned int nSamples, unsigned int nClasses, const float *weights,
                  unsigned int seed, int device)
    : ImageSampler<type, nChannels>(nSamples, seed), m_device(device)
  {
    if (weights) m_weights.assign(weights, weights + nClasses);
    m_desc.kind = kind;
    m_desc.n_samples = nSamples;
    m_desc.seed = seed;
    m_desc.n_classes = nClasses;
    m_desc.weights = NULL;
  }
  pd_sampler desc() const
  {
    pd_sampler s = m_desc;
    s.weights = m_weights.empty() ? NULL : m_weights.data();
    return s;
  }

public:
  unsigned int capacity() const
  {
    const pd_sampler s = desc();
    uint32_t cap = 0;
    if (pd_sampler_capacity(&s, &cap) < 0) throw std::runtime_error(std::string("padenti_hip: ") + pd_last_error());
    return cap;
  }
  unsigned int sample(const type *data, const unsigned char *labels, unsigned int width, unsigned int height,
                      unsigned int *samples) const
  {
    std::vector<std::vector<unsigned int> > out;
    sampleBatch(1, &data, &labels, &width, &height, out);
    std::copy(out[0].begin(), out[0].end(), samples);
    return (unsigned int)out[0].size();
  }
  void sampleBatch(size_t n, const type *const *data, const unsigned char *const *labels, const unsigned int *widths,
                   const unsigned int *heights, std::vector<std::vector<unsigned int> > &out) const
  {
    pd_cxx::sample_on_device<type, nChannels>(m_device, desc(), n, data, labels, widths, heights, out);
  }
};

template <typename type, unsigned int nChannels>
class UniformImageSampler : public GPUImageSampler<type, nChannels>
{
public:
  UniformImageSampler(unsigned int nSamples, unsigned int seed = 0, int device = 0)
    : GPUImageSampler<type, nChannels>(PD_SAMPLER_UNIFORM, nSamples, 0, NULL, seed, device)
  {
  }
};

template <typename type, unsigned int nChannels>
class ClassWeightedImageSampler : public GPUImageSampler<type, nChannels>
{
public:
  ClassWeightedImageSampler(unsigned int nSamples, unsigned int nClasses, unsigned int seed = 0, int device = 0)
    : GPUImageSampler<type, nChannels>(PD_SAMPLER_CLASS_WEIGHTED, nSamples, nClasses, NULL, seed, device)
  {
  }
  ClassWeightedImageSampler(unsigned int nSamples, unsigned int nClasses, const float *weights, unsigned int seed = 0,
                            int device = 0)
    : GPUImageSampler<type, nChannels>(PD_SAMPLER_CLASS_WEIGHTED, nSamples, nClasses, weights, seed, device)
  {
  }
};
