// HIPClassifier: the drop-in for CLClassifier (src/padenti/cl_classifier.hpp:
// 81-94, cl_classifier_impl.hpp:146-557) over libpadenti_hip.so.  operator<<
// copies a tree to the device (the tree may be destroyed afterwards).
#pragma once
#include <algorithm>
#include <stdexcept>
#include <string>

#include "../padenti_hip.h"
#include "classifier.hpp"
#include "hip_tree_trainer.hpp"

typedef int (*pd_forest_factory)(pd_forest **out, int device, uint32_t n_classes);

template <typename ImgType, unsigned int nChannels, typename FeatType, unsigned int FeatDim, unsigned int nClasses>
class HIPClassifier : public Classifier<ImgType, nChannels, FeatType, FeatDim, nClasses>
{
private:
  pd_forest *m_fo;
  HIPClassifier(const HIPClassifier &);
  HIPClassifier &operator=(const HIPClassifier &);

public:
  explicit HIPClassifier(int featureKind, int device = 0) : m_fo(NULL)
  {
    pd_cxx::check(pd_forest_create(&m_fo, device, featureKind, nClasses));
  }
  explicit HIPClassifier(pd_forest_factory create, int device = 0) : m_fo(NULL)
  {
    pd_cxx::check(create(&m_fo, device, nClasses));
  }
  ~HIPClassifier()
  {
    if (m_fo) pd_forest_destroy(m_fo);
  }
  HIPClassifier &operator<<(const Tree<FeatType, FeatDim, nClasses> &tree)
  {
    pd_cxx::check(pd_forest_add_tree(m_fo, tree.getDepth(), tree.getLeftChildren(), tree.getFeatures(),
                                     tree.getThresholds(), tree.getPosteriors()));
    return *this;
  }
  int getNTrees() const { return pd_forest_n_trees(m_fo); }

  void predict(unsigned int id, const Image<ImgType, nChannels> &image, Image<int, 1> &prediction)
  {
    pd_cxx::check(pd_forest_predict_leaves(m_fo, id, image.getData(), image.getWidth(), image.getHeight(), NULL,
                                           prediction.getData()));
  }
  void predict(unsigned int id, const Image<ImgType, nChannels> &image, Image<int, 1> &prediction,
               Image<unsigned char, 1> &mask)
  {
    pd_cxx::check(pd_forest_predict_leaves(m_fo, id, image.getData(), image.getWidth(), image.getHeight(),
                                           mask.getData(), prediction.getData()));
  }
  void predict(const Image<ImgType, nChannels> &image, Image<float, nClasses> &prediction)
  {
    pd_cxx::check(pd_forest_predict_posterior(m_fo, image.getData(), image.getWidth(), image.getHeight(), NULL,
                                              prediction.getData()));
  }
  void predict(const Image<ImgType, nChannels> &image, Image<float, nClasses> &prediction,
               Image<unsigned char, 1> &mask)
  {
    pd_cxx::check(pd_forest_predict_posterior(m_fo, image.getData(), image.getWidth(), image.getHeight(),
                                              mask.getData(), prediction.getData()));
  }
};
