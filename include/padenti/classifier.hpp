// Classifier interface (src/padenti/classifier.hpp:39-96): leaf indices of
// one tree, or the forest's per-class posteriors, optionally masked.
#pragma once
#include "image.hpp"
#include "tree.hpp"

template <typename ImgType, unsigned int nChannels, typename FeatType, unsigned int FeatDim, unsigned int nClasses>
class Classifier
{
public:
  virtual void predict(unsigned int id, const Image<ImgType, nChannels> &image, Image<int, 1> &prediction) = 0;
  virtual void predict(unsigned int id, const Image<ImgType, nChannels> &image, Image<int, 1> &prediction,
                       Image<unsigned char, 1> &mask) = 0;
  virtual void predict(const Image<ImgType, nChannels> &image, Image<float, nClasses> &prediction) = 0;
  virtual void predict(const Image<ImgType, nChannels> &image, Image<float, nClasses> &prediction,
                       Image<unsigned char, 1> &mask) = 0;
  virtual ~Classifier() {}
};
